// mdg_pair.hip -- real-space nonbonded kernels.
//
//  * k_pair_ref64: the reference pair loop (energy.c:480-561 for Ewald,
//    energy.c:216-345 for the cutoff Coulomb functors) over a list in the
//    reference format, FP64, one thread per i in neighOrder order.
//  * k_excl64: the excluded-pair erf correction (energy.c:567-621).
//  * k_pair_fp64 / k_pair_mixed: the same physics over the cluster-pair
//    list, one 64-lane wavefront per chunk (up to 32 j-clusters of one
//    i-cluster), lane (a, b) = pair (i_a, j_b) of each 8x8 cluster pair.
//    i forces stay in registers across the chunk and are reduced once over
//    b; j forces are reduced over a with DPP moves after each cluster pair.
//    k_pair_mixed: FP32 pair arithmetic (displacements formed in FP64
//    first) with int64 fixed-point force accumulation; k_pair_fp64: FP64
//    throughout.
#include "mdg_internal.h"
#include "mdg_side.h"

// ewald.c:356-369
__device__ inline double ewald_dir64(double alpha, double qel, double r, double rt, double &dE)
{
    double ar = alpha * r;
    double e = qel * erfc(ar) * rt;
    double d = 2.0 * qel * alpha * exp(-(ar * ar)) / SQRTPI;
    dE = -(e + d) * rt;
    return e;
}

// vdw.c:144-184
__device__ inline double vdw_switch64(const MdgParams &P, double veps, double vsig, double r2,
                                      double rt, double &dE)
{
    double s = vsig * rt;
    double s3 = s * s * s;
    double s6 = s3 * s3;
    double s12 = s6 * s6;
    if (r2 <= P.cutOn2) {
        dE = 24. * veps * rt * (s6 - 2. * s12);
        return 4. * veps * (s12 - s6);
    }
    double sw1 = P.cutOff2 - r2;
    double S = sw1 * sw1 * (P.cutOff2 + 2. * r2 - 3. * P.cutOn2) * P.switch2;
    double dS = 12. * r2 * sw1 * (P.cutOn2 - r2) * P.switch2;
    double pv = 4. * veps * (s12 - s6);
    double dpv = 24. * veps * (s6 - 2. * s12);
    dE = (pv * dS + dpv * S) * rt;
    return pv * S;
}

// elec.c:137-263: SHIFT1 (2), SHIFT2 (3), SWITCH (4); qel = q_i q_j.
__device__ inline double coulomb64(const MdgParams &P, int type, double qel, double r2, double rt,
                                   double &dE)
{
    double pe = P.chargeConst * qel * rt;
    if (type == 2) {
        double s1 = rt * r2 * P.rcutOff, s2 = r2 * P.rcutOff2;
        double f = 1. - 2. * s1 + s2, df = 2. * (s2 - s1);
        dE = pe * rt * (df - f);
        return pe * f;
    }
    if (type == 3) {
        double s1 = r2 * P.rcutOff2, s2 = s1 * s1;
        double f = 1. - 2. * s1 + s2, df = 4. * (s2 - s1);
        dE = pe * rt * (df - f);
        return pe * f;
    }
    if (type == 4) {
        if (r2 <= P.cutOn2) {
            dE = -pe * rt;
            return pe;
        }
        double sw1 = P.cutOff2 - r2;
        double f = sw1 * sw1 * (P.cutOff2 + 2. * r2 - 3. * P.cutOn2) * P.switch2;
        double df = 12. * r2 * sw1 * (P.cutOn2 - r2) * P.switch2;
        dE = pe * rt * (df - f);
        return pe * f;
    }
    dE = 0.;
    return 0.;
}

// ---------------------------------------------------------------- reference list

// energies out per block: [eElec, eVdw, virElec, virVdw, nIn, nListed]
template <int ELEC>
__global__ void __launch_bounds__(256)
k_pair_ref64(MdgParams P, MdgBox B, const double *__restrict__ x, const double *__restrict__ y,
             const double *__restrict__ z, const double *__restrict__ q,
             const double *__restrict__ eps, const double *__restrict__ sig,
             const int *__restrict__ neighList, const long *__restrict__ off,
             const int *__restrict__ order, double *fx, double *fy, double *fz,
             double *__restrict__ part)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (l < P.n) {
        const int i = order[l];
        const double xi = x[i], yi = y[i], zi = z[i], qi = q[i], ei = eps[i], si = sig[i];
        double fxi = 0., fyi = 0., fzi = 0.;
        for (long k = off[l]; k < off[l + 1]; k++) {
            const int j = neighList[k];
            double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
            double r2 = mdg_dist(B, dx, dy, dz);
            acc[5] += 1.;
            if (r2 <= P.cutOff2) {
                double r = sqrt(r2), rt = 1. / r;
                double dEl, dVw, el;
                if (ELEC == 0) el = ewald_dir64(P.alpha, P.chargeConst * qi * q[j], r, rt, dEl);
                else el = coulomb64(P, ELEC, qi * q[j], r2, rt, dEl);
                double ev = 0.;
                dVw = 0.;
                if (P.vdwOn) ev = vdw_switch64(P, ei * eps[j], si + sig[j], r2, rt, dVw);
                acc[0] += el;
                acc[1] += ev;
                acc[2] += dEl * r;
                acc[3] += dVw * r;
                acc[4] += 1.;
                double fs = (dEl + dVw);
                double fxj = fs * dx * rt, fyj = fs * dy * rt, fzj = fs * dz * rt;
                fxi += fxj;
                fyi += fyj;
                fzi += fzj;
                atomicAdd(&fx[j], -fxj);
                atomicAdd(&fy[j], -fyj);
                atomicAdd(&fz[j], -fzj);
            }
        }
        atomicAdd(&fx[i], fxi);
        atomicAdd(&fy[i], fyi);
        atomicAdd(&fz[i], fzi);
    }
    mdg_block_partials<6>(acc, part);
}

// energy.c:567-621 -- per block: [eCorr, virCorr]
__global__ void __launch_bounds__(256)
k_excl64(MdgParams P, MdgBox B, const double *__restrict__ x, const double *__restrict__ y,
         const double *__restrict__ z, const double *__restrict__ q, const int *__restrict__ exOff,
         const int *__restrict__ exLst, double *fx, double *fy, double *fz,
         double *__restrict__ part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0, 0};
    if (i < P.n) {
        double fxi = 0., fyi = 0., fzi = 0.;
        for (int k = exOff[i]; k < exOff[i + 1]; k++) {
            int j = exLst[k];
            if (j <= i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = mdg_dist(B, dx, dy, dz);
            double r = sqrt(r2), rt = 1. / r, dC;
            acc[0] += ewald_corr64(P.alpha, P.chargeConst * q[i] * q[j], r, rt, dC);
            acc[1] += dC * r;
            double fxj = dC * dx * rt, fyj = dC * dy * rt, fzj = dC * dz * rt;
            fxi += fxj;
            fyi += fyj;
            fzi += fzj;
            atomicAdd(&fx[j], -fxj);
            atomicAdd(&fy[j], -fyj);
            atomicAdd(&fz[j], -fzj);
        }
        if (fxi != 0. || fyi != 0. || fzi != 0.) {
            atomicAdd(&fx[i], fxi);
            atomicAdd(&fy[i], fyi);
            atomicAdd(&fz[i], fzi);
        }
    }
    mdg_block_partials<2>(acc, part);
}

int mdg_pair_ref64_launch(hipStream_t stream, const MdgParams &P, const MdgBox &B, int elecType,
                          const double *x, const double *y, const double *z, const double *q,
                          const double *eps, const double *sig, const int *neighList,
                          const long *off, const int *order, double *fx, double *fy, double *fz,
                          double *part)
{
    const int nb = nblk(P.n, 256);
#define MDG_REF_LAUNCH(E)                                                                          \
    hipLaunchKernelGGL(k_pair_ref64<E>, dim3(nb), dim3(256), 0, stream, P, B, x, y, z, q, eps,    \
                       sig, neighList, off, order, fx, fy, fz, part)
    switch (elecType) {
    case 0: MDG_REF_LAUNCH(0); break;
    case 1: MDG_REF_LAUNCH(1); break;
    case 2: MDG_REF_LAUNCH(2); break;
    case 3: MDG_REF_LAUNCH(3); break;
    case 4: MDG_REF_LAUNCH(4); break;
    default: return MDG_EUNSUPPORTED;
    }
#undef MDG_REF_LAUNCH
    return hipGetLastError() == hipSuccess ? MDG_OK : MDG_EHIP;
}

int mdg_pair_ref64(mdg_ctx *ctx, int elecType, double *fx, double *fy, double *fz,
                   const int *neighList, const long *off, const int *order, double *eout)
{
    const int n = ctx->n, nb = nblk(n, 256);
    double *part = ctx->epart + PART_REF;
    const int st = mdg_pair_ref64_launch(ctx->stream, ctx->P, ctx->B, elecType, ctx->x, ctx->y,
                                         ctx->z, ctx->q, ctx->eps, ctx->sig, neighList, off, order,
                                         fx, fy, fz, part);
    if (st == MDG_EUNSUPPORTED) ctx->err = "unsupported elecType";
    if (st == MDG_EHIP) ctx->err = "kernel launch failed (k_pair_ref64)";
    RETCHK(st);
    RETCHK(mdg_reduce_partials(ctx, part, nb, 6, eout));
    return MDG_OK;
}

int mdg_excl_corr64(mdg_ctx *ctx, double *fx, double *fy, double *fz, double *eout)
{
    const int n = ctx->n, nb = nblk(n, 256);
    double *part = ctx->epart + PART_EXCL;
    hipLaunchKernelGGL(k_excl64, dim3(nb), dim3(256), 0, ctx->stream, ctx->P, ctx->B, ctx->x,
                       ctx->y, ctx->z, ctx->q, ctx->exOff, ctx->exLst, fx, fy, fz, part);
    LAUNCHCHK();
    RETCHK(mdg_reduce_partials(ctx, part, nb, 2, eout));
    return MDG_OK;
}

// ---------------------------------------------------------------- cluster kernels

// FP32 constants of the mixed pair kernel.
struct PairF {
    float alpha, c2a, qc, cutOn2, cutOff2, switch2, threeOn2;
    float swA, swB, swD;             // 2 switch2, (cutOff2 - 3 cutOn2) switch2, 12 switch2
    float rcutOff, rcutOff2;         // 1/rc, 1/rc^2 (SHIFT1 / SHIFT2)
    float rsDelta;                   // relative coordinates: exact test within this of cutOff2
    double cutOff2d, L0, L1, L2, iL0, iL1, iL2;
};

// The electrostatics of a cluster pair kernel: the Ewald real-space term
// (energy.c:496-503, keyEwald 2), none (elec noelec / full evaluate
// nothing, energy.c:45-48), or the cutoff functors SHIFT1 / SHIFT2 /
// SWITCH of elec.c:137-263 (keyEwald 0, energy.c:216-345).
enum { PAIR_EWALD = 0, PAIR_NONE = 1, PAIR_SHIFT1 = 2, PAIR_SHIFT2 = 3, PAIR_SWITCH = 4 };

// erfc(x) for x >= 0 as t Q(t) exp(-x^2), t = 1 / (1 + x / 2).  The
// substitution is that of the Numerical Recipes "erfcc" form (Press et al.,
// 2nd ed., sec. 6.2), which writes erfc as t exp(P(t) - x^2); here
// Q(t) = erfcx(x) / t itself is the polynomial, so one exponential serves
// both erfc and the force term exp(-x^2) (returned in *ez).  Q has degree 9,
// fitted by mdbas_amd/erfc_fit.py (weighted least squares for the relative
// error over t in [0, 1], i.e. every x >= 0): relative error of the fit
// 1.6e-7; in FP32 about 2e-6 at x = 6, from the rounding of x^2 inside the
// exponential.
__device__ inline float mdg_erfc_q(float x, float *ez)
{
    const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));
    float p = fmaf(t, -0.02552288340259605f, 0.06446153649429703f);
    p = fmaf(t, p, 0.030753194918781607f);
    p = fmaf(t, p, -0.1985488152101896f);
    p = fmaf(t, p, 0.07946913878476664f);
    p = fmaf(t, p, 0.05826160866113146f);
    p = fmaf(t, p, 0.18041489004864147f);
    p = fmaf(t, p, 0.24651220482179143f);
    p = fmaf(t, p, 0.28210421927957413f);
    p = fmaf(t, p, 0.28209474704728354f);
    const float e = __expf(-(x * x));
    *ez = e;
    return t * p * e;
}

struct ClusterArgs {
    MdgParams P;
    MdgBox B;
    PairF F;
    int nclu;
    const double *sx, *sy, *sz;      // slot positions
    const float4 *sp32;              // slot (q, eps, sig, 0) for MIXED
    const double *sq, *seps, *ssig;  // slot parameters for FP64
    const int *cpOff, *cpJ;
    const unsigned long long *cpMask;
    const int *chOff;                // per i-cluster chunk offsets (nclu + 1)
    int maxChunks;                   // chunk capacity (an overflowing build is flagged)
    int cap;                         // entry capacity of cpJ / cpMask
    const int2 *chunk;               // chunk -> (I, first entry)
    long long *fix, *fiy, *fiz;      // MIXED: int64 fixed-point slot forces
    double *fdx, *fdy, *fdz;         // FP64: slot forces
    double *part;                    // per block: eDir, eVdw, virDir, virVdw, nIn
    const MdgBox *boxes;             // batch: box of each replica
    const int *crep;                 // batch: replica of each cluster
    double *chE;                     // batch: per chunk eDir, eVdw, virDir, virVdw, nIn
    long fstr;                       // fiy - fix and fiz - fiy (one allocation)
    // pruned list (mixed): per chunk the kept entries at the chunk's offsets
    const int *pJ;
    const unsigned long long *pMask;
    const int *pCnt;
    const double *disp;              // per slot block: max squared move (k_slot_pos)
    int nDisp;
    double pruneLim;                 // pruned list usable while 4 max move^2 <= pruneLim
    int *ctrl;
    // relative coordinates (mixed, one rectangular box; see mx_pair_rel)
    const float4 *spr;               // per slot (null: FP64 displacements only)
    const float4 *cpD, *pD;          // per entry of the full / pruned list
    const double *relPart;           // per slot-writing block: max |relative component|
    int nRel;
    double relLim;
    int relExact;                    // every entry's cutoff test in FP64 (a check mode)
};

// The three j-force components of each aligned group of 8 lanes (the 8
// i atoms of one j atom), summed and scattered in three DPP steps: the
// group's lanes a = 0, 1 end with the x sum, a = 2, 3 with y, a = 4, 5
// with z (a = 6, 7 with zero).  Step 1 pairs lane a with 7 - a
// (row_half_mirror): lanes a < 4 keep x and y and receive the partner's,
// lanes a >= 4 keep z; step 2 pairs a with a ^ 2 (quad_perm [2,3,0,1]),
// each keeping one of its two values; step 3 adds lane a ^ 1's.
template <int CTRL>
__device__ inline float dpp_movf(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

__device__ inline float dpp_rs8(float fx, float fy, float fz, int a)
{
    const bool hi = a >= 4;
    float k0 = hi ? fz : fx, k1 = hi ? 0.f : fy;
    const float s0 = hi ? fx : fz, s1 = hi ? fy : 0.f;
    k0 += dpp_movf<0x141>(s0);
    k1 += dpp_movf<0x141>(s1);
    const bool b1 = (a & 2) != 0;
    float k = b1 ? k1 : k0;
    k += dpp_movf<0x4E>(b1 ? k0 : k1);
    k += dpp_movf<0xB1>(k);
    return k;
}

// Sum over the 8 lanes of each aligned group of 8 with DPP moves (plain
// VALU ops, no LDS): quad_perm [1,0,3,2], quad_perm [2,3,0,1], then
// row_half_mirror (lane i <- lane 7-i within the half row).
__device__ inline double dpp_sum8(double v)
{
    auto mv = [](double x, int ctrl) {
        long long b = __double_as_longlong(x);
        int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
        switch (ctrl) {
        case 0: lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, true);
                hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, true); break;
        case 1: lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, true);
                hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, true); break;
        default: lo = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xF, 0xF, true);
                 hi = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xF, 0xF, true); break;
        }
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    v += mv(v, 0);
    v += mv(v, 1);
    v += mv(v, 2);
    return v;
}

// Lane mapping: a = lane & 7 is the i atom, b = lane >> 3 the j atom, so
// each j-cluster pair's j-force sums run over 8 adjacent lanes (DPP) and
// the i-force sums over lanes a, a+8, ..., a+56 happen once per chunk.
// Mask bit 8 b + a (= the lane) marks the listed pair (I_a, J_b).  FP64
// throughout.
template <bool BATCH, int ELEC>
__global__ void __launch_bounds__(256) k_pair_fp64(ClusterArgs A)
{
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, b = lane >> 3;
    const int bit = lane;
    double e_acc[5] = {0, 0, 0, 0, 0};
    const int nch = min(A.chOff[A.nclu], A.maxChunks);
    for (int W = blockIdx.x * 4 + (threadIdx.x >> 6); W < nch; W += gridDim.x * 4) {
        // chunk W: entries [e0, e1) of i-cluster I (at most MDG_CHUNK of them)
        const int2 ch = A.chunk[W];
        const int I = ch.x, e0 = ch.y, e1 = min(min(e0 + MDG_CHUNK, A.cpOff[I + 1]), A.cap);
        if (e0 >= e1) {   // only after a flagged list overflow
            if (BATCH && lane < 5) A.chE[(long)W * 5 + lane] = 0.;
            continue;
        }
        // the box: the run's, or the replica's in a batch
        const MdgBox &Bx = BATCH ? A.boxes[A.crep[I]] : A.B;
        const int si = I * MDG_CL + a;
        const double xi = A.sx[si], yi = A.sy[si], zi = A.sz[si];
        const double qid = A.sq[si], eid = A.seps[si], gid = A.ssig[si];
        double fxid = 0., fyid = 0., fzid = 0.;
        double ed = 0., ev = 0., vd = 0., vv = 0.;
        int nin = 0;
        for (int e = e0; e < e1; e++) {
            const int J = A.cpJ[e];
            const unsigned long long m = A.cpMask[e];
            const int sj = J * MDG_CL + b;
            double dx = A.sx[sj] - xi, dy = A.sy[sj] - yi, dz = A.sz[sj] - zi;
            const double r2 = mdg_dist(Bx, dx, dy, dz);
            const bool inCut = ((m >> bit) & 1ull) && r2 <= A.P.cutOff2;
            // an entry with no pair inside the cutoff adds only exact zeros
            if (__builtin_amdgcn_ballot_w64(inCut) == 0) continue;
            double fjx = 0., fjy = 0., fjz = 0.;
            if (inCut) {
                const double r = sqrt(r2), rt = 1. / r;
                double dEl = 0., dVw = 0., el = 0., evw = 0.;
                if (ELEC == PAIR_EWALD)
                    el = ewald_dir64(A.P.alpha, A.P.chargeConst * qid * A.sq[sj], r, rt, dEl);
                else if (ELEC != PAIR_NONE)
                    el = coulomb64(A.P, ELEC, qid * A.sq[sj], r2, rt, dEl);
                if (A.P.vdwOn)
                    evw = vdw_switch64(A.P, eid * A.seps[sj], gid + A.ssig[sj], r2, rt, dVw);
                const double fs = dEl + dVw;
                fjx = fs * dx * rt;
                fjy = fs * dy * rt;
                fjz = fs * dz * rt;
                ed += el;
                ev += evw;
                vd += dEl * r;
                vv += dVw * r;
                nin++;
            }
            fxid += fjx;
            fyid += fjy;
            fzid += fjz;
            fjx = dpp_sum8(fjx);
            fjy = dpp_sum8(fjy);
            fjz = dpp_sum8(fjz);
            if (a == 0) {
                if (fjx != 0.) atomicAdd(&A.fdx[sj], -fjx);
                if (fjy != 0.) atomicAdd(&A.fdy[sj], -fjy);
                if (fjz != 0.) atomicAdd(&A.fdz[sj], -fjz);
            }
        }
        for (int o = 8; o < 64; o <<= 1) {
            fxid += __shfl_xor(fxid, o, 64);
            fyid += __shfl_xor(fyid, o, 64);
            fzid += __shfl_xor(fzid, o, 64);
        }
        if (b == 0) {
            atomicAdd(&A.fdx[si], fxid);
            atomicAdd(&A.fdy[si], fyid);
            atomicAdd(&A.fdz[si], fzid);
        }
        if (BATCH) {
            // energies per chunk, summed per replica later (a replica's
            // chunks are contiguous)
            const double c[5] = {mdg_wave_sum(ed), mdg_wave_sum(ev), mdg_wave_sum(vd),
                                 mdg_wave_sum(vv), mdg_wave_sum((double)nin)};
            if (lane < 5) A.chE[(long)W * 5 + lane] = c[lane];
        } else {
            e_acc[0] += ed;
            e_acc[1] += ev;
            e_acc[2] += vd;
            e_acc[3] += vv;
            e_acc[4] += nin;
        }
    }
    if (!BATCH) mdg_block_partials<5>(e_acc, A.part);
}

// Mixed precision over the same list and lane mapping: displacements in
// FP64 from the positions (mx_pair) or in FP32 relative coordinates
// (mx_pair_rel), the cutoff test the reference's FP64 one either way, pair
// arithmetic in FP32, energies summed in FP64 (each FP32 term, before any
// rounding of a partial sum: the total is a sum of ~1e7 terms of +-1e4
// with heavy cancellation), virials in FP32 over a chunk and FP64 across
// chunks, forces added as int64 fixed point (deterministic).
//
// The loop runs on wave-uniform scalars: chunk, entry, j cluster and mask
// come through the scalar cache, and the j slots are read with buffer loads
// whose offset is J * (slot bytes * 8) in a scalar register plus a fixed
// per-lane offset.  Two register sets alternate so the next entry's j data
// is in flight while the current one is computed.  There is no branch per
// pair: the list mask and the cutoff test combine as wave masks (the mask
// bit of a pair is its lane), and lanes outside either evaluate a harmless
// distance with zero coefficients.  The j-force sums of a cluster pair go
// out as one atomic per component: after the reduce-scatter (dpp_rs8)
// lanes a = 0, 2, 4 of each j group add x, y, z.
struct MxJ {
    double x, y, z;
    float4 p;
};

struct MxSlots {
    __amdgpu_buffer_rsrc_t x, y, z, p, r;
    int vo8, vo16;                   // per-lane byte offsets of slot b
};

__device__ __forceinline__ MxJ mx_load(const MxSlots &R, int J)
{
    typedef int i4v __attribute__((ext_vector_type(4)));
    MxJ o;
    o.x = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(R.x, R.vo8, J * 64, 0));
    o.y = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(R.y, R.vo8, J * 64, 0));
    o.z = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(R.z, R.vo8, J * 64, 0));
    const i4v v = __builtin_amdgcn_raw_buffer_load_b128(R.p, R.vo16, J * 128, 0);
    o.p = make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z),
                      __int_as_float(v.w));
    return o;
}

struct MxAcc {
    float fx, fy, fz;                // i-atom force
    double ed, ev;                   // energies, summed per pair in FP64
    float vd, vv;                    // virials: FP32 over the chunk (at most
                                     // MDG_CHUNK terms per lane), FP64 across chunks
    int nin;                         // in-cutoff pairs (wave-uniform)
};

// floor(v * 2^32) as int64 for |v| < 2^31: the integer part floor(v) and
// the fraction v - floor(v) are both exact in FP32 (the fraction keeps a
// subset of v's significand bits), and so is the fraction times 2^32
// (< 2^32), whose truncation gives the low word.
__device__ inline long long mdg_f32_fix(float v)
{
    const float fl = floorf(v);
    const unsigned int lo = (unsigned int)((v - fl) * 4294967296.0f);
    return (long long)(((unsigned long long)(unsigned int)(int)fl << 32) | lo);
}

// The reference's minimum-image r^2 of one pair (utils.c:38-80, 76): the
// shift L * n is exact for |n| <= 1 (slot positions are never more than a
// box edge apart) and so is the difference when n != 0 (Sterbenz), so
// rint's ties-to-even and the reference's nint (utils.c:606-609) give the
// same displacement wherever |dx| < L / 2 - rounding, i.e. for every pair
// that can be inside the cutoff; each product is rounded, so the cutoff
// test decides as energy.c:500 does.
__device__ inline double mx_ref_r2(double &dx, double &dy, double &dz, double L0, double L1,
                                   double L2, double iL0, double iL1, double iL2)
{
#pragma clang fp contract(off)
    dx = dx - L0 * rint(dx * iL0);
    dy = dy - L1 * rint(dy * iL1);
    dz = dz - L2 * rint(dz * iL2);
    return (dx * dx + dy * dy) + dz * dz;
}

// The pair arithmetic of one entry once its in-cutoff lanes inm are known:
// (dx, dy, dz) = x_j - x_i and r2f its square (FP32), r2 the reference's
// FP64 r^2 (read only by the cutoff functors' FP64 energy).
template <bool BATCH, int ELEC>
__device__ __forceinline__ void mx_body(const ClusterArgs &A, const float4 &jp,
                                        unsigned long long inm, int J, float dxf, float dyf,
                                        float dzf, float r2in, double r2, float qi, float ei,
                                        float gi, double qid, int a, int b, MxAcc &acc)
{
    const bool in = __builtin_amdgcn_inverse_ballot_w64(inm);
    acc.nin += __builtin_popcountll(inm);
    const float r2f = in ? r2in : A.F.cutOff2;
    const float qq = in ? qi * jp.x : 0.f;
    const float veps = (in && A.P.vdwOn) ? ei * jp.y : 0.f;
    const float rinv = __builtin_amdgcn_rsqf(r2f);
    const float r = r2f * rinv;
    const float s = (gi + jp.z) * rinv;
    const float s2 = s * s;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float pv = 4.f * veps * (s12 - s6);
    const float dpv = 24.f * veps * (s6 - 2.f * s12);
    // switch of vdw.c:144-184 with r2 clamped from below at cutOn2, where
    // dS vanishes exactly; S is 1 below cutOn2
    const bool sw = r2f > A.F.cutOn2;
    const float u = sw ? r2f : A.F.cutOn2;
    const float sw1 = A.F.cutOff2 - u;
    // S = sw1^2 (cutOff2 + 2 u - 3 cutOn2) switch2 and
    // dS = 12 u sw1 (cutOn2 - u) switch2, the constants folded on the host
    const float Su = (sw1 * sw1) * __builtin_fmaf(u, A.F.swA, A.F.swB);
    const float S = sw ? Su : 1.f;
    const float dS = (u * sw1) * (A.F.cutOn2 - u) * A.F.swD;
    const float evdw = pv * S;
    const float dvdw = (pv * dS + dpv * S) * rinv;
    float eel, del;
    double eel64 = 0.;   // the cutoff functors' pair energy, see below
    if (ELEC == PAIR_EWALD) {
        float ez;
        const float erfcv = mdg_erfc_q(A.F.alpha * r, &ez);
        eel = qq * erfcv * rinv;
        del = -(eel + qq * A.F.c2a * ez) * rinv;
    } else if (ELEC == PAIR_NONE) {
        eel = del = 0.f;
    } else {
        // pe * f and pe / r * (f' r - f), elec.c:137-263; SWITCH shares
        // the LJ switch S (1 below cutOn2) and its r dS/dr.  The shifted
        // forms are evaluated as squares, f = (1 - s)^2 with s = r / rc
        // (SHIFT1) or r^2 / rc^2 (SHIFT2), the same polynomials as the
        // reference's 1 - 2 s + s^2 without its cancellation near rc, where
        // FP32 would lose the small f
        const float pe = qq * rinv;
        float f, df;
        if (ELEC == PAIR_SHIFT1) {
            const float s1 = r * A.F.rcutOff, t = 1.f - s1;
            f = t * t;
            df = -2.f * s1 * t;       // 2 (s^2 - s)
        } else if (ELEC == PAIR_SHIFT2) {
            const float s1 = r2f * A.F.rcutOff2, t = 1.f - s1;
            f = t * t;
            df = -4.f * s1 * t;       // 4 (s^2 - s)
        } else {
            f = S;
            df = dS;
        }
        eel = pe * f;
        del = pe * rinv * (df - f);
        // The unscreened pair energies sum ~1e9 in magnitude to a total
        // ~1e3 times smaller, and FP32 terms carry fixed rounding biases
        // (the charge product of each atom-type pair, the approximate
        // reciprocal square root), so the energy term is formed in FP64
        // from the FP64 r^2 (the force keeps FP32)
        if (in) {
            double ri = (double)rinv;
            ri = ri * (1.5 - 0.5 * r2 * ri * ri);          // Newton step: ~1e-14
            ri = ri * (1.5 - 0.5 * r2 * ri * ri);
            double f64;
            if (ELEC == PAIR_SHIFT1) {
                const double t = 1. - r2 * ri * A.P.rcutOff;
                f64 = t * t;
            } else if (ELEC == PAIR_SHIFT2) {
                const double t = 1. - r2 * A.P.rcutOff2;
                f64 = t * t;
            } else if (r2 <= A.P.cutOn2) {
                f64 = 1.;
            } else {
                const double sw1 = A.P.cutOff2 - r2;
                f64 = sw1 * sw1 * (A.P.cutOff2 + 2. * r2 - 3. * A.P.cutOn2) * A.P.switch2;
            }
            eel64 = (qid * A.sq[J * MDG_CL + b]) * ri * f64;
        }
    }
    const float fs = (del + dvdw) * rinv;
    const float fjx = fs * dxf, fjy = fs * dyf, fjz = fs * dzf;
    if (ELEC == PAIR_EWALD || ELEC == PAIR_NONE) acc.ed += eel;
    else acc.ed += eel64;
    acc.ev += evdw;
    acc.vd += del * r;
    acc.vv += dvdw * r;
    acc.fx += fjx;
    acc.fy += fjy;
    acc.fz += fjz;
    // lanes a = 0, 2, 4 of each j group add x, y, z
    const float gc = dpp_rs8(fjx, fjy, fjz, a);
    if ((a & 1) == 0 && a < 6) {
        const long long iv = mdg_f32_fix(-gc);
        if (iv != 0)
            atomicAdd((unsigned long long *)(A.fix + (a >> 1) * A.fstr + J * MDG_CL + b),
                      (unsigned long long)iv);
    }
}

// One entry with the displacement in FP64 from the slot positions.
template <bool BATCH, int ELEC>
__device__ __forceinline__ void mx_pair(const ClusterArgs &A, const MxJ &j, unsigned long long m,
                                        int J, double xi, double yi, double zi, float qi, float ei,
                                        float gi, double qid, double L0, double L1, double L2,
                                        double iL0, double iL1, double iL2, int a, int b,
                                        MxAcc &acc)
{
    double dx, dy, dz, r2;
    {
#pragma clang fp contract(off)
        dx = j.x - xi;
        dy = j.y - yi;
        dz = j.z - zi;
    }
    r2 = mx_ref_r2(dx, dy, dz, L0, L1, L2, iL0, iL1, iL2);
    const unsigned long long inm = m & __builtin_amdgcn_ballot_w64(r2 <= A.F.cutOff2d);
    // an entry with no pair inside the cutoff adds exact zeros everywhere
    // (zero coefficients, no j-force atomics): skip it, wave-uniformly.
    // The list keeps entries with a pair within cutOff + delr (or + the
    // prune skin), so a sixth to a third of them have none.
    if (inm == 0) return;
    mx_body<BATCH, ELEC>(A, j.p, inm, J, (float)dx, (float)dy, (float)dz, (float)r2, r2, qi, ei,
                         gi, qid, a, b, acc);
}

// One entry in relative coordinates (one rectangular box, Ewald or no
// electrostatics): x_j - x_i = D + (r_j - r_i) in FP32, with D the
// entry's minimum image of O_J - O_I and r the slots' minimum images from
// their cluster origins.  While every |r| component stays below relLim
// ((min edge / 2 - cutoff) / 2 less a margin), this is the minimum image
// of every pair inside the cutoff (any other image is at least an edge
// minus the cutoff away).  Its FP32 r^2 is within rsDelta / 2 of the
// reference's FP64 r^2 there (the bound is formed on the host from the
// FP32 roundings), so lanes farther than rsDelta from the cutoff decide as
// the reference does; an entry with a listed lane closer than that (or
// every entry under relExact) takes the FP64 test of mx_pair for its
// mask.  The results are the reference's in-cutoff set exactly.
struct MxR {
    float4 r, p;
};

template <int ELEC>
__device__ __forceinline__ void mx_pair_rel(const ClusterArgs &A, const MxR &j,
                                            unsigned long long m, float4 D, int J, float4 ri,
                                            int si, float qi, float ei, float gi, int a, int b,
                                            MxAcc &acc)
{
    const float dx = D.x + (j.r.x - ri.x), dy = D.y + (j.r.y - ri.y), dz = D.z + (j.r.z - ri.z);
    const float r2f = dx * dx + dy * dy + dz * dz;
    const unsigned long long near =
        m & __builtin_amdgcn_ballot_w64(fabsf(r2f - A.F.cutOff2) < A.F.rsDelta);
    unsigned long long inm;
    if (near == 0 && !A.relExact) {
        inm = m & __builtin_amdgcn_ballot_w64(r2f <= A.F.cutOff2);
    } else {
        const int sj = J * MDG_CL + b;
        double ex, ey, ez;
        {
#pragma clang fp contract(off)
            ex = A.sx[sj] - A.sx[si];
            ey = A.sy[sj] - A.sy[si];
            ez = A.sz[sj] - A.sz[si];
        }
        const double r2 = mx_ref_r2(ex, ey, ez, A.F.L0, A.F.L1, A.F.L2, A.F.iL0, A.F.iL1, A.F.iL2);
        inm = m & __builtin_amdgcn_ballot_w64(r2 <= A.F.cutOff2d);
    }
    if (inm == 0) return;
    mx_body<false, ELEC>(A, j.p, inm, J, dx, dy, dz, r2f, 0., qi, ei, gi, 0., a, b, acc);
}

// A wave-uniform float4 through the scalar cache (two 64-bit loads).
__device__ inline float4 mx_uload_f4(const float4 *p, long i)
{
    const unsigned long long *q = (const unsigned long long *)p;
    const unsigned long long lo = mdg_uload(q, 2 * i), hi = mdg_uload(q, 2 * i + 1);
    return make_float4(__uint_as_float((unsigned int)lo), __uint_as_float((unsigned int)(lo >> 32)),
                       __uint_as_float((unsigned int)hi), __uint_as_float((unsigned int)(hi >> 32)));
}

__device__ __forceinline__ MxR mx_load_rel(const MxSlots &R, int J)
{
    typedef int i4v __attribute__((ext_vector_type(4)));
    MxR o;
    const i4v u = __builtin_amdgcn_raw_buffer_load_b128(R.r, R.vo16, J * 128, 0);
    o.r = make_float4(__int_as_float(u.x), __int_as_float(u.y), __int_as_float(u.z),
                      __int_as_float(u.w));
    const i4v v = __builtin_amdgcn_raw_buffer_load_b128(R.p, R.vo16, J * 128, 0);
    o.p = make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z),
                      __int_as_float(v.w));
    return o;
}

// Whether this launch runs in relative coordinates: the context has them,
// the electrostatics allow it, and the slot writer's per-block largest
// |component| stays within relLim.
template <bool BATCH, int ELEC>
__device__ inline bool mx_use_rel(const ClusterArgs &A, int lane)
{
    if (BATCH || !(ELEC == PAIR_EWALD || ELEC == PAIR_NONE) || !A.spr || A.nRel <= 0) return false;
    double r = 0.;
    for (int k = lane; k < A.nRel; k += 64) r = fmax(r, A.relPart[k]);
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    return __builtin_amdgcn_readfirstlane((int)(r <= A.relLim)) != 0;
}

// REL (one box, Ewald or no electrostatics, a box the relative coordinates
// fit): the relative-coordinate loop, with a plain FP64-displacement loop
// for launches whose guard fails; otherwise the FP64-displacement loop with
// its loads double-buffered.  The two variants differ in register use (the
// relative loop alone fits six wavefronts per SIMD).
template <bool BATCH, int ELEC, bool REL>
__global__ void __launch_bounds__(256) k_pair_mixed(ClusterArgs A)
{
    // ahead of the SPME kernels beside it when the launch shape asks for it
    if (A.P.spmePrio == 2) __builtin_amdgcn_s_setprio(2);
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, b = lane >> 3;
    const long nsb = (long)A.nclu * MDG_CL;
    MxSlots R;
    R.x = __builtin_amdgcn_make_buffer_rsrc((void *)A.sx, 0, (int)(nsb * 8), 0x00020000);
    R.y = __builtin_amdgcn_make_buffer_rsrc((void *)A.sy, 0, (int)(nsb * 8), 0x00020000);
    R.z = __builtin_amdgcn_make_buffer_rsrc((void *)A.sz, 0, (int)(nsb * 8), 0x00020000);
    R.p = __builtin_amdgcn_make_buffer_rsrc((void *)A.sp32, 0, (int)(nsb * 16), 0x00020000);
    R.r = __builtin_amdgcn_make_buffer_rsrc((void *)A.spr, 0, A.spr ? (int)(nsb * 16) : 0,
                                            0x00020000);
    R.vo8 = b * 8;
    R.vo16 = b * 16;
    double e_acc[5] = {0, 0, 0, 0, 0};
    const int nch = min(mdg_uload(A.chOff, A.nclu), A.maxChunks);
    const int w0 = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    // the pruned list when no atom moved more than skin / 2 since it was made
    // (the entries it drops contribute exact zeros, so the results are the
    // same bit for bit)
    bool pruned = false;
    if (A.pruneLim > 0.) {
        double d2 = 0.;
        for (int k = lane; k < A.nDisp; k += 64) d2 = fmax(d2, A.disp[k]);
        for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
        pruned = __builtin_amdgcn_readfirstlane((int)(4. * d2 <= A.pruneLim)) != 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.ctrl[CTRL_PRUNE] = 0;
        if (pruned) A.ctrl[CTRL_PRUNED] += 1;
    }
    const int *__restrict__ eJ = pruned ? A.pJ : A.cpJ;
    const unsigned long long *__restrict__ eM = pruned ? A.pMask : A.cpMask;
    const bool rel = REL && mx_use_rel<BATCH, ELEC>(A, lane);
    const float4 *__restrict__ eD = pruned ? A.pD : A.cpD;
    for (int W = w0; W < nch; W += gridDim.x * 4) {
        const long long ch = mdg_uload((const long long *)A.chunk, W);   // int2 (I, e0)
        const int I = (int)(ch & 0xffffffffll), e0 = (int)(ch >> 32);
        const int e1 = pruned ? min(e0 + mdg_uload(A.pCnt, W), A.cap)
                              : min(min(e0 + MDG_CHUNK, mdg_uload(A.cpOff, I + 1)), A.cap);
        if (e0 >= e1) {   // an empty pruned chunk, or a flagged list overflow
            if (BATCH && lane < 5) A.chE[(long)W * 5 + lane] = 0.;
            continue;
        }
        double L0 = A.F.L0, L1 = A.F.L1, L2 = A.F.L2, iL0 = A.F.iL0, iL1 = A.F.iL1, iL2 = A.F.iL2;
        if (BATCH) {
            const MdgBox &Bx = A.boxes[mdg_uload(A.crep, I)];
            L0 = Bx.a1;
            L1 = Bx.type == 1 ? Bx.a1 : Bx.b2;
            L2 = Bx.type == 1 ? Bx.a1 : Bx.c3;
            iL0 = Bx.u1;
            iL1 = Bx.type == 1 ? Bx.u1 : Bx.v2;
            iL2 = Bx.type == 1 ? Bx.u1 : Bx.w3;
        }
        const int si = I * MDG_CL + a;
        const float4 pi = A.sp32[si];
        const float qi = A.F.qc * pi.x;
        MxAcc acc = {0.f, 0.f, 0.f, 0., 0., 0.f, 0.f, 0};
        // entries e (set 0) and e + 1 (set 1) alternate; indices past the
        // end are clamped to the last entry (loaded, never computed)
        if (REL && rel) {
            const float4 ri = A.spr[si];
            int J0 = mdg_uload(eJ, e0);
            unsigned long long m0 = mdg_uload(eM, e0);
            float4 D0 = mx_uload_f4(eD, e0);
            MxR j0 = mx_load_rel(R, J0);
            for (int e = e0;;) {
                const int J1 = mdg_uload(eJ, min(e + 1, e1 - 1));
                const unsigned long long m1 = mdg_uload(eM, min(e + 1, e1 - 1));
                const float4 D1 = mx_uload_f4(eD, min(e + 1, e1 - 1));
                const MxR j1 = mx_load_rel(R, J1);
                mx_pair_rel<ELEC>(A, j0, m0, D0, J0, ri, si, qi, pi.y, pi.z, a, b, acc);
                if (++e >= e1) break;
                J0 = mdg_uload(eJ, min(e + 1, e1 - 1));
                m0 = mdg_uload(eM, min(e + 1, e1 - 1));
                D0 = mx_uload_f4(eD, min(e + 1, e1 - 1));
                j0 = mx_load_rel(R, J0);
                mx_pair_rel<ELEC>(A, j1, m1, D1, J1, ri, si, qi, pi.y, pi.z, a, b, acc);
                if (++e >= e1) break;
            }
        } else if (REL) {
            // the guard failed: FP64 displacements, one entry at a time
            const double xi = A.sx[si], yi = A.sy[si], zi = A.sz[si];
            for (int e = e0; e < e1; e++) {
                const int J = mdg_uload(eJ, e);
                mx_pair<BATCH, ELEC>(A, mx_load(R, J), mdg_uload(eM, e), J, xi, yi, zi, qi, pi.y,
                                     pi.z, 0., L0, L1, L2, iL0, iL1, iL2, a, b, acc);
            }
        } else {
            const double xi = A.sx[si], yi = A.sy[si], zi = A.sz[si];
            const double qid = ELEC >= PAIR_SHIFT1 ? A.P.chargeConst * A.sq[si] : 0.;
            int J0 = mdg_uload(eJ, e0);
            unsigned long long m0 = mdg_uload(eM, e0);
            MxJ j0 = mx_load(R, J0);
            for (int e = e0;;) {
                const int J1 = mdg_uload(eJ, min(e + 1, e1 - 1));
                const unsigned long long m1 = mdg_uload(eM, min(e + 1, e1 - 1));
                const MxJ j1 = mx_load(R, J1);
                mx_pair<BATCH, ELEC>(A, j0, m0, J0, xi, yi, zi, qi, pi.y, pi.z, qid, L0, L1, L2,
                                     iL0, iL1, iL2, a, b, acc);
                if (++e >= e1) break;
                J0 = mdg_uload(eJ, min(e + 1, e1 - 1));
                m0 = mdg_uload(eM, min(e + 1, e1 - 1));
                j0 = mx_load(R, J0);
                mx_pair<BATCH, ELEC>(A, j1, m1, J1, xi, yi, zi, qi, pi.y, pi.z, qid, L0, L1, L2,
                                     iL0, iL1, iL2, a, b, acc);
                if (++e >= e1) break;
            }
        }
        // i forces: sums over the 8 j lanes of each i atom; lanes b = 0, 1,
        // 2 of the group add x, y, z
        double gx = acc.fx, gy = acc.fy, gz = acc.fz;
        for (int o = 8; o < 64; o <<= 1) {
            gx += __shfl_xor(gx, o, 64);
            gy += __shfl_xor(gy, o, 64);
            gz += __shfl_xor(gz, o, 64);
        }
        if (b < 3) mdg_atomic_add_fix(A.fix + b * A.fstr + si, b == 0 ? gx : (b == 1 ? gy : gz),
                                      MDG_FSCALE);
        if (BATCH) {
            const double c[5] = {mdg_wave_sum(acc.ed), mdg_wave_sum(acc.ev),
                                 mdg_wave_sum((double)acc.vd), mdg_wave_sum((double)acc.vv),
                                 (double)acc.nin};
            if (lane < 5) A.chE[(long)W * 5 + lane] = c[lane];
        } else {
            e_acc[0] += acc.ed;
            e_acc[1] += acc.ev;
            e_acc[2] += acc.vd;
            e_acc[3] += acc.vv;
            e_acc[4] += lane == 0 ? (double)acc.nin : 0.;
        }
    }
    if (!BATCH) mdg_block_partials<5>(e_acc, A.part);
}

// Slot-ordered positions for the pair kernels (padding slots get copies
// of a far point; their mask bits are never set), and zeroed slot force
// accumulators of the active precision.
//
// Mixed precision also measures how far the slots moved since the
// positions the pruned list was made from: that list holds, per chunk, the
// entries with a pair within cutOff + skin at those positions, so while no
// atom has moved more than skin / 2 since, every pair it leaves out is
// still beyond the cutoff.  The previous slot positions are those
// positions exactly when the CTRL_PRUNE word is set (the prune kernel sets
// it after reading them, every pair kernel clears it, and the FP64 variant
// of this kernel, which rewrites the slots untracked, clears it too).  Each
// block writes its largest squared move (minimum image), or +inf when the
// pruned list does not belong to the previous positions.
//
// With spr given (mixed, one rectangular box) every variant also writes the
// slots' relative coordinates and per block their largest |component|.
struct SlotMove {
    int *ctrl;
    int fresh;                       // the list was rebuilt: slots changed atoms
    int track;                       // pruning on (mixed, rectangular boxes)
    MdgBox B;
    const MdgBox *boxes;             // batch: box of each replica
    const int *crep;                 // batch: replica of each cluster
    double *part;                    // per block: max squared move
    float4 *spr;                     // relative coordinates (or null)
    const double *cO;                // cluster origins [3][nclu]
    int nclu;
    double *relPart;                 // per block: max |relative component|
};

template <bool MIXED>
__global__ void __launch_bounds__(256)
k_slot_pos(int nslot, const int *__restrict__ slotAtom, const double *__restrict__ x,
           const double *__restrict__ y, const double *__restrict__ z, double *__restrict__ sx,
           double *__restrict__ sy, double *__restrict__ sz, long long *fix, long long *fiy,
           long long *fiz, double *fdx, double *fdy, double *fdz, SlotMove M)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double rel = 0.;
    if (M.spr && s < nslot) {
        const int a = slotAtom[s];
        if (a >= 0) rel = mdg_slot_rel(M.B, M.cO, M.nclu, s / MDG_CL, x[a], y[a], z[a], M.spr, s);
        else M.spr[s] = make_float4(1e6f, 1e6f, 1e6f, 0.f);   // never in a mask
    }
    if (M.spr) mdg_block_max(rel, M.relPart);
    if (!MIXED) {
        if (s == 0) M.ctrl[CTRL_PRUNE] = 0;
        if (s >= nslot) return;
        const int a = slotAtom[s];
        sx[s] = a >= 0 ? x[a] : 1e7;
        sy[s] = a >= 0 ? y[a] : 1e7;
        sz[s] = a >= 0 ? z[a] : 1e7;
        fdx[s] = 0.;
        fdy[s] = 0.;
        fdz[s] = 0.;
        return;
    }
    const bool track = M.track && !M.fresh && M.ctrl[CTRL_PRUNE] == 1;
    double d2 = 0.;
    if (s < nslot) {
        const int a = slotAtom[s];
        const double nx = a >= 0 ? x[a] : 1e7, ny = a >= 0 ? y[a] : 1e7,
                     nz = a >= 0 ? z[a] : 1e7;
        if (track && a >= 0) {
            const MdgBox &Bx = M.boxes ? M.boxes[M.crep[s / MDG_CL]] : M.B;
            const double L0 = Bx.a1, L1 = Bx.type == 1 ? Bx.a1 : Bx.b2,
                         L2 = Bx.type == 1 ? Bx.a1 : Bx.c3;
            double dx = nx - sx[s], dy = ny - sy[s], dz = nz - sz[s];
            dx -= L0 * rint(dx / L0);
            dy -= L1 * rint(dy / L1);
            dz -= L2 * rint(dz / L2);
            d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 <= 1e300)) d2 = INFINITY;   // NaN positions: no pruning
        }
        sx[s] = nx;
        sy[s] = ny;
        sz[s] = nz;
        fix[s] = 0;
        fiy[s] = 0;
        fiz[s] = 0;
    }
    if (!M.part) return;
    for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
    __shared__ double wmax[4];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0)
        M.part[blockIdx.x] =
            track ? fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3])) : INFINITY;
}

__global__ void k_slot_param(int nslot, const int *__restrict__ slotAtom,
                             const double *__restrict__ q, const double *__restrict__ eps,
                             const double *__restrict__ sig, float4 *__restrict__ sp32,
                             double *__restrict__ sq, double *__restrict__ seps,
                             double *__restrict__ ssig)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    int a = slotAtom[s];
    double qa = a >= 0 ? q[a] : 0., ea = a >= 0 ? eps[a] : 0., ga = a >= 0 ? sig[a] : 0.;
    sp32[s] = make_float4((float)qa, (float)ea, (float)ga, 0.f);
    sq[s] = qa;
    seps[s] = ea;
    ssig[s] = ga;
}

int mdg_slot_params(mdg_ctx *ctx)
{
    ClusterList &C = ctx->cl;
    const int nslot = C.nclu * MDG_CL;
    hipLaunchKernelGGL(k_slot_param, dim3(nblk(nslot, 256)), dim3(256), 0, ctx->stream, nslot,
                       C.slotAtom, ctx->q, ctx->eps, ctx->sig, ctx->sp32, ctx->sq, ctx->seps,
                       ctx->ssig);
    LAUNCHCHK();
    return MDG_OK;
}

static ClusterArgs cluster_args(mdg_ctx *ctx)
{
    ClusterList &C = ctx->cl;
    ClusterArgs A;
    A.P = ctx->P;
    A.P.spmePrio = ctx->tune.spmePrio;
    A.B = ctx->B;
    A.F.alpha = (float)ctx->P.alpha;
    A.F.c2a = (float)(2.0 * ctx->P.alpha / SQRTPI);
    A.F.qc = (float)ctx->P.chargeConst;
    A.F.cutOn2 = (float)ctx->P.cutOn2;
    A.F.cutOff2 = (float)ctx->P.cutOff2;
    A.F.switch2 = (float)ctx->P.switch2;
    A.F.threeOn2 = (float)(3.0 * ctx->P.cutOn2);
    A.F.swA = (float)(2.0 * ctx->P.switch2);
    A.F.swB = (float)((ctx->P.cutOff2 - 3.0 * ctx->P.cutOn2) * ctx->P.switch2);
    A.F.swD = (float)(12.0 * ctx->P.switch2);
    A.F.rcutOff = (float)ctx->P.rcutOff;
    A.F.rcutOff2 = (float)ctx->P.rcutOff2;
    A.F.cutOff2d = ctx->P.cutOff2;
    A.F.L0 = ctx->B.a1;
    A.F.L1 = ctx->B.type == 1 ? ctx->B.a1 : ctx->B.b2;
    A.F.L2 = ctx->B.type == 1 ? ctx->B.a1 : ctx->B.c3;
    // the reference's reciprocal edges (u1, v2, w3 of init_box), as dist
    // multiplies by them (utils.c:44-54)
    A.F.iL0 = ctx->B.u1;
    A.F.iL1 = ctx->B.type == 1 ? ctx->B.u1 : ctx->B.v2;
    A.F.iL2 = ctx->B.type == 1 ? ctx->B.u1 : ctx->B.w3;
    A.nclu = C.nclu;
    A.sx = ctx->sx;
    A.sy = ctx->sy;
    A.sz = ctx->sz;
    A.sp32 = ctx->sp32;
    A.sq = ctx->sq;
    A.seps = ctx->seps;
    A.ssig = ctx->ssig;
    A.cpOff = C.cpOff;
    A.cpJ = C.cpJ;
    A.cpMask = C.cpMask;
    A.chOff = C.chOff;
    A.chunk = C.chunk;
    A.maxChunks = C.maxChunks;
    A.cap = (int)C.cap;
    A.fix = ctx->fix;
    A.fiy = ctx->fiy;
    A.fiz = ctx->fiz;
    A.fstr = ctx->fiy - ctx->fix;
    A.fdx = ctx->fdx;
    A.fdy = ctx->fdy;
    A.fdz = ctx->fdz;
    A.part = ctx->epart + PART_PAIR;
    A.boxes = ctx->bt ? ctx->bt->boxes : nullptr;
    A.crep = ctx->bt ? ctx->bt->crep : nullptr;
    A.chE = ctx->bt ? ctx->bt->chE : nullptr;
    A.pJ = C.pJ;
    A.pMask = C.pMask;
    A.pCnt = C.pCnt;
    A.disp = ctx->dispPart;
    A.nDisp = ctx->dispPart ? ctx->nDispPart : 0;
    // usable while 2 max move <= skin, with a relative margin for rounding
    A.pruneLim = ctx->pruneSkin > 0. && C.pJ ? ctx->pruneSkin * ctx->pruneSkin * (1. - 1e-6) : -1.;
    A.ctrl = ctx->ctrl;
    // relative coordinates (the launch sets nRel from the slot writer)
    A.spr = nullptr;
    A.cpD = C.cpD;
    A.pD = C.pD;
    A.relPart = ctx->relPart;
    A.nRel = 0;
    A.relLim = -1.;
    A.relExact = ctx->pairPath == 2 ? 1 : 0;
    A.F.rsDelta = 0.f;
    if (ctx->precision == MDG_MIXED && !ctx->bt && C.cO && C.cpD && ctx->spr &&
        ctx->pairPath != 1) {
        const double L[3] = {A.F.L0, A.F.L1, A.F.L2};
        const double lmin = std::min(L[0], std::min(L[1], L[2]));
        const double lmax = std::max(L[0], std::max(L[1], L[2]));
        const double rc = ctx->P.cutOff, rc2 = ctx->P.cutOff2;
        // every |relative component| below relLim keeps 2 relLim + rc below
        // half the shortest edge, with 0.1 A to spare
        const double lim = 0.5 * (0.5 * lmin - rc) - 0.05;
        if (lim > 0.) {
            // FP32 unit in the last place of v > 0
            auto ulp = [](double v) { return std::ldexp(1.0, std::ilogb(v) - 23); };
            // per component: two relative coordinates (one rounding each),
            // D (|D| <= half an edge), their difference and the final sum
            // near the cutoff, each rounded to nearest
            const double eps = ulp(lim) + 0.5 * ulp(0.5 * lmax) + 0.5 * ulp(2. * lim) +
                               0.5 * ulp(1.01 * rc);
            // |r2f - r^2| near the cutoff: 2 |d|_1 eps + 3 eps^2, the three
            // roundings of the sum of squares, the FP32 cutoff's own rounding
            const double bound = 2. * std::sqrt(3.) * 1.01 * rc * eps + 3. * eps * eps +
                                 1.5 * ulp(1.01 * rc2) + 0.5 * ulp(rc2) + 1e-12 * rc2;
            A.spr = ctx->spr;
            A.relLim = lim;
            A.F.rsDelta = (float)(4. * bound);
        }
    }
    return A;
}

// Slot positions (k_slot_pos) of the active precision.
static int slot_pos(mdg_ctx *ctx, bool mixed)
{
    ClusterList &C = ctx->cl;
    const int nslot = C.nclu * MDG_CL;
    SlotMove M;
    M.ctrl = ctx->ctrl;
    M.fresh = ctx->pruneFresh ? 1 : 0;
    M.track = ctx->pruneSkin > 0. ? 1 : 0;
    M.B = ctx->B;
    M.boxes = ctx->bt ? ctx->bt->boxes : nullptr;
    M.crep = ctx->bt ? ctx->bt->crep : nullptr;
    M.part = ctx->dispPart;
    // relative coordinates whenever the list has cluster origins (every
    // variant keeps them in step with the slot positions)
    M.spr = C.cO && ctx->spr ? ctx->spr : nullptr;
    M.cO = C.cO;
    M.nclu = C.nclu;
    M.relPart = ctx->relPart;
    if (mixed) {
        ctx->pruneFresh = false;
        hipLaunchKernelGGL(k_slot_pos<true>, dim3(nblk(nslot, 256)), dim3(256), 0, ctx->stream,
                           nslot, C.slotAtom, ctx->x, ctx->y, ctx->z, ctx->sx, ctx->sy, ctx->sz,
                           ctx->fix, ctx->fiy, ctx->fiz, ctx->fdx, ctx->fdy, ctx->fdz, M);
    } else {
        hipLaunchKernelGGL(k_slot_pos<false>, dim3(nblk(nslot, 256)), dim3(256), 0, ctx->stream,
                           nslot, C.slotAtom, ctx->x, ctx->y, ctx->z, ctx->sx, ctx->sy, ctx->sz,
                           ctx->fix, ctx->fiy, ctx->fiz, ctx->fdx, ctx->fdy, ctx->fdz, M);
    }
    LAUNCHCHK();
    return MDG_OK;
}

// The kernel variant of the context's electrostatics.
static int pair_elec(const mdg_ctx *ctx)
{
    if (ctx->P.ewald != 0) return PAIR_EWALD;
    const int t = ctx->P.elecType;
    return (t >= PAIR_SHIFT1 && t <= PAIR_SWITCH) ? t : PAIR_NONE;
}

template <bool BATCH>
static void launch_pair(mdg_ctx *ctx, bool mixed, int nb, const ClusterArgs &A)
{
#define MDG_PAIR_LAUNCH(E)                                                                         \
    do {                                                                                           \
        if (mixed && !BATCH && A.spr && (E == PAIR_EWALD || E == PAIR_NONE))                       \
            hipLaunchKernelGGL((k_pair_mixed<false, E, true>), dim3(nb), dim3(256), 0, ctx->stream, \
                               A);                                                                 \
        else if (mixed)                                                                            \
            hipLaunchKernelGGL((k_pair_mixed<BATCH, E, false>), dim3(nb), dim3(256), 0,            \
                               ctx->stream, A);                                                    \
        else                                                                                       \
            hipLaunchKernelGGL((k_pair_fp64<BATCH, E>), dim3(nb), dim3(256), 0, ctx->stream, A);   \
    } while (0)
    switch (pair_elec(ctx)) {
    case PAIR_EWALD: MDG_PAIR_LAUNCH(PAIR_EWALD); break;
    case PAIR_SHIFT1: MDG_PAIR_LAUNCH(PAIR_SHIFT1); break;
    case PAIR_SHIFT2: MDG_PAIR_LAUNCH(PAIR_SHIFT2); break;
    case PAIR_SWITCH: MDG_PAIR_LAUNCH(PAIR_SWITCH); break;
    default: MDG_PAIR_LAUNCH(PAIR_NONE); break;
    }
#undef MDG_PAIR_LAUNCH
}

int mdg_pair_cluster(mdg_ctx *ctx)
{
    const bool mixed = ctx->precision == MDG_MIXED;
    int nDisp = ctx->nDispPart;
    if (mixed && ctx->slotsDone)   // this step's stage 1 wrote the slots
        nDisp = mdg_vv_blocks(ctx);
    else
        RETCHK(slot_pos(ctx, mixed));
    ctx->slotsDone = false;
    ClusterArgs A = cluster_args(ctx);
    if (A.nDisp) A.nDisp = nDisp;
    if (A.spr) A.nRel = nDisp;   // the slot writer just before wrote the relative coordinates
    const int nb = mdg_pair_blocks(ctx);
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[0], ctx->stream));
    launch_pair<false>(ctx, mixed, nb, A);
    LAUNCHCHK();
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[1], ctx->stream));
    return MDG_OK;
}

// The list prune as a kernel of its own (prune_waves, mdg_side.h).
template <bool BATCH>
__global__ void __launch_bounds__(SIDE_BLOCK) k_prune(PruneArgs A)
{
    prune_waves<BATCH>(A, blockIdx.x * (SIDE_BLOCK / 64) + (threadIdx.x >> 6),
                       gridDim.x * (SIDE_BLOCK / 64));
}

PruneArgs mdg_prune_args(mdg_ctx *ctx)
{
    ClusterList &C = ctx->cl;
    const ClusterArgs A = cluster_args(ctx);
    PruneArgs P;
    P.nclu = C.nclu;
    P.maxChunks = C.maxChunks;
    P.cap = (int)C.cap;
    P.chOff = C.chOff;
    P.cpOff = C.cpOff;
    P.cpJ = C.cpJ;
    P.cpMask = C.cpMask;
    P.chunk = C.chunk;
    P.sx = ctx->sx;
    P.sy = ctx->sy;
    P.sz = ctx->sz;
    P.L0 = A.F.L0;
    P.L1 = A.F.L1;
    P.L2 = A.F.L2;
    P.iL0 = A.F.iL0;
    P.iL1 = A.F.iL1;
    P.iL2 = A.F.iL2;
    P.boxes = A.boxes;
    P.crep = A.crep;
    P.pJ = C.pJ;
    P.pMask32 = (unsigned int *)C.pMask;
    P.pCnt = C.pCnt;
    P.cpD = C.cpD;
    P.pD = C.cpD ? C.pD : nullptr;
    P.rp = ctx->P.cutOff + ctx->pruneSkin;
    P.rp2 = P.rp * P.rp;
    P.ctrl = ctx->ctrl;
    return P;
}

bool mdg_prune_on(const mdg_ctx *ctx)
{
    return ctx->precision == MDG_MIXED && ctx->pruneSkin > 0. && ctx->cl.pJ && ctx->dispPart;
}

int mdg_prune_blocks(const mdg_ctx *ctx)
{
    return std::min(nblk(ctx->cl.maxChunks, SIDE_BLOCK / 64), MDG_PAIR_MAX_BLOCKS);
}

int mdg_prune_list(mdg_ctx *ctx)
{
    if (!mdg_prune_on(ctx)) return MDG_OK;
    const PruneArgs P = mdg_prune_args(ctx);
    const int nb = mdg_prune_blocks(ctx);
    if (ctx->bt)
        hipLaunchKernelGGL((k_prune<true>), dim3(nb), dim3(SIDE_BLOCK), 0, ctx->stream, P);
    else
        hipLaunchKernelGGL((k_prune<false>), dim3(nb), dim3(SIDE_BLOCK), 0, ctx->stream, P);
    LAUNCHCHK();
    return MDG_OK;
}

// Adjoint of the real-space pair energy with respect to the per-atom
// parameters, FP64 in both precisions.  For every listed pair inside the
// cutoff (the pairs energy.c:496-541 sums):
//   dE/dq_i   += C q_j erfc(a r) / r                      (and i <-> j)
//   dE/deps_i += eps_j 4 (s^12 - s^6) S(r)                (and i <-> j)
//   dE/dsig_i += 4 eps_i eps_j (12 s^11 - 6 s^5) S(r) / r (both atoms)
// with s = (sig_i + sig_j) / r and S the switch of vdw.c:144-184 (1 below
// cutOn).  Same wavefront layout as k_pair_fp64; slot sums go to
// fdx (q), fdy (eps), fdz (sig).
__global__ void __launch_bounds__(256) k_adj_cluster(ClusterArgs A)
{
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, b = lane >> 3;
    const int bit = lane;
    const int nch = min(A.chOff[A.nclu], A.maxChunks);
    for (int W = blockIdx.x * 4 + (threadIdx.x >> 6); W < nch; W += gridDim.x * 4) {
        const int2 ch = A.chunk[W];
        const int I = ch.x, e0 = ch.y, e1 = min(min(e0 + MDG_CHUNK, A.cpOff[I + 1]), A.cap);
        const int si = I * MDG_CL + a;
        const double xi = A.sx[si], yi = A.sy[si], zi = A.sz[si];
        const double qi = A.sq[si], ei = A.seps[si], gi = A.ssig[si];
        double gqi = 0., gei = 0., gsi = 0.;
        for (int e = e0; e < e1; e++) {
            const int J = A.cpJ[e];
            const unsigned long long m = A.cpMask[e];
            const int sj = J * MDG_CL + b;
            double dx = A.sx[sj] - xi, dy = A.sy[sj] - yi, dz = A.sz[sj] - zi;
            const double r2 = mdg_dist(A.B, dx, dy, dz);
            double gqj = 0., gej = 0., gsj = 0.;
            if (((m >> bit) & 1ull) && r2 <= A.P.cutOff2) {
                const double r = sqrt(r2), rt = 1. / r;
                const double qj = A.sq[sj], ej = A.seps[sj];
                const double ec = A.P.chargeConst * erfc(A.P.alpha * r) * rt;
                gqi += qj * ec;
                gqj = qi * ec;
                const double s = (gi + A.ssig[sj]) * rt, s2 = s * s, s6 = s2 * s2 * s2;
                const double s12 = s6 * s6, s5 = s2 * s2 * s, s11 = s5 * s6;
                double S = 1.;
                if (r2 > A.P.cutOn2) {
                    const double sw = A.P.cutOff2 - r2;
                    S = sw * sw * (A.P.cutOff2 + 2. * r2 - 3. * A.P.cutOn2) * A.P.switch2;
                }
                const double lj = 4. * (s12 - s6) * S;
                gei += ej * lj;
                gej = ei * lj;
                const double gs = 4. * ei * ej * (12. * s11 - 6. * s5) * rt * S;
                gsi += gs;
                gsj = gs;
            }
            gqj = dpp_sum8(gqj);
            gej = dpp_sum8(gej);
            gsj = dpp_sum8(gsj);
            if (a == 0) {
                if (gqj != 0.) atomicAdd(&A.fdx[sj], gqj);
                if (gej != 0.) atomicAdd(&A.fdy[sj], gej);
                if (gsj != 0.) atomicAdd(&A.fdz[sj], gsj);
            }
        }
        for (int o = 8; o < 64; o <<= 1) {
            gqi += __shfl_xor(gqi, o, 64);
            gei += __shfl_xor(gei, o, 64);
            gsi += __shfl_xor(gsi, o, 64);
        }
        if (b == 0) {
            atomicAdd(&A.fdx[si], gqi);
            atomicAdd(&A.fdy[si], gei);
            atomicAdd(&A.fdz[si], gsi);
        }
    }
}

// Real-space parameter gradients into the FP64 slot accumulators.
int mdg_adj_pair(mdg_ctx *ctx)
{
    RETCHK(slot_pos(ctx, false));
    ClusterArgs A = cluster_args(ctx);
    hipLaunchKernelGGL(k_adj_cluster, dim3(mdg_pair_blocks(ctx)), dim3(256), 0, ctx->stream, A);
    LAUNCHCHK();
    return MDG_OK;
}

// Radial derivatives per unit veps of the switched LJ of vdw.c:144-184,
// V = pv S, pv = 4 (s^12 - s^6), s = vsig / r: d1 = V', d2 = V'',
// ds = dV'/dvsig.
__device__ inline void hvp_lj(const MdgParams &P, double vsig, double r, double &d1, double &d2,
                              double &ds)
{
    const double r2 = r * r, rt = 1. / r, s = vsig * rt;
    const double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6, s5 = s2 * s2 * s, s11 = s6 * s5;
    const double pv = 4. * (s12 - s6);
    const double pv1 = 24. * (s6 - 2. * s12) * rt;
    const double pv2 = 4. * (156. * s12 - 42. * s6) * rt * rt;
    const double pvs = 4. * (12. * s11 - 6. * s5) * rt;
    const double pv1s = 24. * (6. * s5 - 24. * s11) * rt * rt;
    double S = 1., S1 = 0., S2 = 0.;
    if (r2 > P.cutOn2) {
        const double sw1 = P.cutOff2 - r2, on = P.cutOn2 - r2;
        S = sw1 * sw1 * (P.cutOff2 + 2. * r2 - 3. * P.cutOn2) * P.switch2;
        S1 = 12. * r * sw1 * on * P.switch2;
        S2 = 12. * P.switch2 * (sw1 * on - 2. * r2 * on - 2. * r2 * sw1);
    }
    d1 = pv1 * S + pv * S1;
    d2 = pv2 * S + 2. * pv1 * S1 + pv * S2;
    ds = pv1s * S + pvs * S1;
}

// Second derivatives of the real-space pair energy over the device list
// (the pairs of k_pair_fp64, r^2 <= cutOff^2), for a direction u in slot
// order: with d = x_j - x_i, phi the pair energy, g = phi'/r and
// h = (phi'' - g)/r^2,
//   (H u)_j = -(H u)_i = g du + h (d . du) d,   du = u_j - u_i,
// and the pair's share of d/dtheta (grad E . u) = d/dtheta g (d . du) for
// theta = q_i, q_j, eps_i, eps_j, sig_i, sig_j.  Same wavefront layout as
// k_adj_cluster; FP64, slot sums with atomics.
__global__ void __launch_bounds__(256) k_hvp_cluster(ClusterArgs A, HvpSlots H)
{
    const int lane = threadIdx.x & 63;
    const int a = lane & 7, b = lane >> 3;
    const int bit = lane;
    const int nch = min(A.chOff[A.nclu], A.maxChunks);
    const double C = A.P.chargeConst, al = A.P.alpha;
    for (int W = blockIdx.x * 4 + (threadIdx.x >> 6); W < nch; W += gridDim.x * 4) {
        const int2 ch = A.chunk[W];
        const int I = ch.x, e0 = ch.y, e1 = min(min(e0 + MDG_CHUNK, A.cpOff[I + 1]), A.cap);
        const int si = I * MDG_CL + a;
        const double xi = A.sx[si], yi = A.sy[si], zi = A.sz[si];
        const double qi = A.sq[si], ei = A.seps[si], gi = A.ssig[si];
        const double uxi = H.ux[si], uyi = H.uy[si], uzi = H.uz[si];
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int e = e0; e < e1; e++) {
            const int J = A.cpJ[e];
            const unsigned long long m = A.cpMask[e];
            const int sj = J * MDG_CL + b;
            double dx = A.sx[sj] - xi, dy = A.sy[sj] - yi, dz = A.sz[sj] - zi;
            const double r2 = mdg_dist(A.B, dx, dy, dz);
            double w[6] = {0, 0, 0, 0, 0, 0};
            if (((m >> bit) & 1ull) && r2 <= A.P.cutOff2) {
                const double r = sqrt(r2), rt = 1. / r, ar = al * r;
                const double qj = A.sq[sj], ej = A.seps[sj];
                const double ec = erfc(ar), gs = 2. * al * exp(-ar * ar) / SQRTPI;
                const double el1 = -(ec * rt + gs) * rt;                      // per unit qel
                const double el2 = 2. * ec * rt * rt * rt + gs * (2. * rt * rt + 2. * al * al);
                double l1, l2, ls;
                hvp_lj(A.P, gi + A.ssig[sj], r, l1, l2, ls);
                const double qel = C * qi * qj, veps = ei * ej;
                const double d1 = qel * el1 + veps * l1, d2 = qel * el2 + veps * l2;
                const double g = d1 * rt, h = (d2 - g) * rt * rt;
                const double dux = H.ux[sj] - uxi, duy = H.uy[sj] - uyi, duz = H.uz[sj] - uzi;
                const double t = dx * dux + dy * duy + dz * duz;
                w[0] = g * dux + h * t * dx;
                w[1] = g * duy + h * t * dy;
                w[2] = g * duz + h * t * dz;
                const double cq = C * el1 * rt * t, cl = l1 * rt * t, cs = veps * ls * rt * t;
                acc[0] -= w[0];
                acc[1] -= w[1];
                acc[2] -= w[2];
                acc[3] += qj * cq;
                acc[4] += ej * cl;
                acc[5] += cs;
                w[3] = qi * cq;
                w[4] = ei * cl;
                w[5] = cs;
            }
            double *out[6] = {H.hx, H.hy, H.hz, H.mq, H.me, H.ms};
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const double v = dpp_sum8(w[k]);
                if (a == 0 && v != 0.) atomicAdd(&out[k][sj], v);
            }
        }
        double *out[6] = {H.hx, H.hy, H.hz, H.mq, H.me, H.ms};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double v = acc[k];
            for (int o = 8; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
            if (b == 0) atomicAdd(&out[k][si], v);
        }
    }
}

// Real-space H u and mixed terms into the slot accumulators of H (which
// the caller zeroes), at the context's positions, over the current list.
int mdg_hvp_pair(mdg_ctx *ctx, const HvpSlots &H)
{
    RETCHK(slot_pos(ctx, false));
    ClusterArgs A = cluster_args(ctx);
    hipLaunchKernelGGL(k_hvp_cluster, dim3(mdg_pair_blocks(ctx)), dim3(256), 0, ctx->stream, A, H);
    LAUNCHCHK();
    return MDG_OK;
}

// Pair forces of every replica of a batch: slot positions (and zeroed
// accumulators), then the cluster kernel with per-replica boxes and
// per-chunk energies.
int mdg_b_pair(mdg_ctx *ctx)
{
    const bool mixed = ctx->precision == MDG_MIXED;
    RETCHK(slot_pos(ctx, mixed));
    ClusterArgs A = cluster_args(ctx);
    const int nb = mdg_pair_blocks(ctx);
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[0], ctx->stream));
    launch_pair<true>(ctx, mixed, nb, A);
    LAUNCHCHK();
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[1], ctx->stream));
    return MDG_OK;
}
