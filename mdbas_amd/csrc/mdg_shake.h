// mdg_shake.h -- shared state of the SHAKE kernels (generic groups in
// mdg_integrate.hip, three-site molecules in mdg_shake_w3.hip).
#pragma once
#include "mdg_internal.h"
#include "mdg_side.h"

// Window of SHAKE states kept per group around its first convergence, and
// a ring of SHK_R for the last sweeps at which a whole wavefront was
// converged.
#define SHK_W 8
#define SHK_R 4
// [0] start, [1..W] window, [W+1] end, [W+2 .. W+1+R] ring
#define SHK_SLOTS (SHK_W + 2 + SHK_R)
#define SHK_MASKW 4               // 256 sweep indices

struct ShakeArgs {
    MdgParams P;
    MdgBox B;
    int nGrp;
    const int *atomOff, *atoms, *conOff, *cons;
    const int *ca, *cb;
    const double *rc2, *rmass, *mass;
    const double *thr;    // per constraint: the stage-1 test 0.5 |r^2 - d^2| / d < tol
                          // holds exactly when |r^2 - d^2| < thr (d = sqrt(rc2))
    const int *nAtConst;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *fx, *fy, *fz;
    double *dd;           // [3][nConst]
    int nConst;
    int *cycle;
    double *vir;
    int *ctrl;
    double *part;         // stage 2: kin, stresk[6]
    double *scr;          // [SHK_SLOTS][6][n] saved states
    double *virScr;       // SHAKE virial at those states: [SHK_SLOTS][nGrp] for generic
                          // groups, [SHK_SLOTS][nGrp][3] (per constraint) for three-site
    int *first, *kend;    // per group: first converged sweep, last checked sweep
    int *ring;            // three-site kernels, per group: the sweep in each ring
                          // slot (SHK_R ints, -1 when empty)
    unsigned long long *mask;   // device-wide AND of converged-sweep bits
    // replica batch: box and mask per replica, replica per group, and the
    // per-group sum m v^2 (stage 2)
    const MdgBox *boxes;
    const int *grep;
    double *grpKin;
    // three-site kernels: one molecule per quad of lanes, on every
    // spread-th quad of a wavefront (16 / spread molecules per wavefront);
    // 0: one molecule per lane (the k_m3_* kernels)
    int spread;
    // stage 1 of a device step without a list rebuild (one box, mixed):
    // k_w3_vv1_b also writes the slot positions and clears the slot force
    // accumulators, as k_slot_pos<true> would, with the per-block largest
    // squared move into dispPart (see k_slot_pos)
    int slotWrite, track;
    const int *atomSlot;
    double *sx, *sy, *sz;
    long long *fix;
    long fstr;
    double *dispPart;
    // and the slots' relative coordinates when spr is given (mdg_slot_rel),
    // with per block their largest |component| into relPart
    float4 *spr;
    const double *cO;
    int ncluO;
    double *relPart;
    // thermostat and barostat integrators (generic-group kernels only):
    // ddStored takes the constraint vectors from dd, formed at the call's
    // entry, instead of from the positions the kernel loads; noKick skips
    // the half kick; npt drifts as x = dscale (x - com) + dt v + com; noWrap
    // leaves stage 2's image_update to a later kernel
    int ddStored = 0, noKick = 0, npt = 0, noWrap = 0;
    double dscale = 1., com0 = 0., com1 = 0., com2 = 0.;
};

// Molecule of this lane's quad, or nGrp for a quad left idle.
__device__ inline int shake_w3_group(const ShakeArgs &A)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int qd = lane >> 2;
    if (qd % A.spread) return A.nGrp;
    const int g = wave * (16 / A.spread) + qd / A.spread;
    return g < A.nGrp ? g : A.nGrp;
}

// Bits of word w that count as converged without a check: sweeps past
// kend of a group that was converged at its last two checks, and every
// sweep index from maxCycle on (the stopping rule never reads those, so a
// word that lies wholly beyond the last sweep that matters is all ones).
__device__ inline unsigned long long shake_fill_bits(int w, int kend, bool settled, int maxCycle)
{
    const int lo = (settled ? min(kend + 1, maxCycle) : maxCycle) - w * 64;
    if (lo <= 0) return ~0ull;
    if (lo >= 64) return 0ull;
    return ~0ull << lo;
}

// Publishes the group's converged-at-sweep-k bits into the device-wide
// AND mask: bits for k <= kend are as checked; later sweeps count as
// converged when the group was converged at its last two checks.
//
// Atomics on one address serialise (about 11 ns each on MI355X), so the
// bits are ANDed within the block first and each block issues one atomic
// per word that is not all ones.  Every thread of the block must call this.
__device__ inline void shake_publish_mask(const ShakeArgs &A, const unsigned long long (&bits)[SHK_MASKW],
                                          int kend, bool settled)
{
    __shared__ unsigned long long wm[SHK_MASKW][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int w = 0; w < SHK_MASKW; w++) {
        unsigned long long m = bits[w] | shake_fill_bits(w, kend, settled, A.P.maxCycle);
        for (int o = 32; o > 0; o >>= 1) m &= __shfl_xor(m, o, 64);
        if (lane == 0) wm[w][wid] = m;
    }
    __syncthreads();
    if (threadIdx.x < SHK_MASKW) {
        const int nw = (blockDim.x + 63) >> 6;
        unsigned long long m = ~0ull;
        for (int v = 0; v < nw; v++) m &= wm[threadIdx.x][v];
        if (m != ~0ull) atomicAnd(&A.mask[threadIdx.x], m);
    }
}

// Batch form of shake_publish_mask: each replica has its own mask
// (masks + SHK_MASKW * rep); a wavefront whose groups all belong to one
// replica publishes once, a wavefront on a replica boundary lane by lane.
// Idle lanes pass rep = -1.
__device__ inline void shake_publish_mask_b(const ShakeArgs &A,
                                            const unsigned long long (&bits)[SHK_MASKW], int kend,
                                            bool settled, int rep)
{
    const int lane = threadIdx.x & 63;
    int lo = rep < 0 ? 0x7fffffff : rep, hi = rep;
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    const bool uniform = lo == hi;
    for (int w = 0; w < SHK_MASKW; w++) {
        unsigned long long m = bits[w] | shake_fill_bits(w, kend, settled, A.P.maxCycle);
        if (uniform) {
            for (int o = 32; o > 0; o >>= 1) m &= __shfl_xor(m, o, 64);
            if (lane == 0 && hi >= 0 && m != ~0ull) atomicAnd(&A.mask[(long)hi * SHK_MASKW + w], m);
        } else if (rep >= 0 && m != ~0ull) {
            atomicAnd(&A.mask[(long)rep * SHK_MASKW + w], m);
        }
    }
}

// First sweep index at which every group was converged (-1 if none); mask
// is the run's (or one replica's) SHK_MASKW words.
__device__ inline int shake_global_stop(const ShakeArgs &A, const unsigned long long *mask)
{
    for (int w = 0; w < SHK_MASKW; w++) {
        unsigned long long m = mask[w];
        if (m) {
            int k = w * 64 + __ffsll((long long)m) - 1;
            return k < A.P.maxCycle ? k : -1;
        }
    }
    return -1;
}

__device__ inline int shake_global_stop(const ShakeArgs &A) { return shake_global_stop(A, A.mask); }


// Which saved state leads to sweep C: a window or ring slot, the end
// state, or the start (when C lies outside both and before the end).  ring:
// the group's SHK_R ring sweeps (-1 when empty).
__device__ inline void shake_pick_slot(int C, int first, int kend, const int *ring, int &slot, int &at)
{
    int rs = -1;
    if (ring)
        for (int q = 0; q < SHK_R; q++)
            if (ring[q] == C) rs = q;
    if (first >= 0 && C >= first && C - first < SHK_W && C <= kend) {
        slot = 1 + C - first;
        at = C;
    } else if (rs >= 0) {
        slot = SHK_W + 2 + rs;
        at = C;
    } else if (C >= kend) {
        slot = SHK_W + 1;
        at = kend;
    } else {
        slot = 0;
        at = 0;
    }
}

// utils.c:162-218 (image_update) for one atom.
__device__ inline void shake_image(const MdgBox &B, double &xs, double &ys, double &zs)
{
#pragma clang fp contract(off)
    if (B.type == 1) {
        xs = xs - B.a1 * mdg_nint(xs * B.u1);
        ys = ys - B.a1 * mdg_nint(ys * B.u1);
        zs = zs - B.a1 * mdg_nint(zs * B.u1);
    } else if (B.type == 2) {
        xs = xs - B.a1 * mdg_nint(xs * B.u1);
        ys = ys - B.b2 * mdg_nint(ys * B.v2);
        zs = zs - B.c3 * mdg_nint(zs * B.w3);
    } else if (B.type == 3) {
        double xt = xs * B.u1 + ys * B.u2 + zs * B.u3;
        double yt = xs * B.v1 + ys * B.v2 + zs * B.v3;
        double zt = xs * B.w1 + ys * B.w2 + zs * B.w3;
        xt -= mdg_nint(xt);
        yt -= mdg_nint(yt);
        zt -= mdg_nint(zt);
        xs = xt * B.a1 + yt * B.b1 + zt * B.c1;
        ys = xt * B.a2 + yt * B.b2 + zt * B.c2;
        zs = xt * B.a3 + yt * B.b3 + zt * B.c3;
    }
}

template <bool BATCH> __global__ void k_w3_vv1_a(ShakeArgs A, SideWork S);
template <bool BATCH> __global__ void k_w3_vv1_b(ShakeArgs A);
template <bool BATCH> __global__ void k_w3_vv2_a(ShakeArgs A, FinArgs F, int fuse);
template <bool BATCH> __global__ void k_w3_vv2_b(ShakeArgs A);
// one molecule per lane (SHAKE spread 0, mdg_shake_m3.hip)
template <bool BATCH> __global__ void k_m3_vv1_a(ShakeArgs A, SideWork S);
template <bool BATCH> __global__ void k_m3_vv1_b(ShakeArgs A);
template <bool BATCH> __global__ void k_m3_vv2_a(ShakeArgs A, FinArgs F, int fuse);
template <bool BATCH> __global__ void k_m3_vv2_b(ShakeArgs A);
