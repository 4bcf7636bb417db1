// mdg_shake_w3.hip -- VV + SHAKE kernels specialised for rigid
// three-site molecules (three atoms a0 < a1 < a2 with constraints
// (a0,a1), (a0,a2), (a1,a2) in that order, e.g. TIP3P water under
// consth).  Same arithmetic, operation order and global stopping rule as
// the generic kernels of mdg_integrate.hip.
//
// One molecule per quad of lanes: lane k < 3 holds atom k's state and
// constraint k's data, lane 3 idles.  Molecules sit on every spread-th
// quad, so a smaller spread packs more molecules per wavefront and a larger
// one gives the SIMDs more wavefronts to interleave (the sweeps are
// latency-bound chains); the autotuner picks it.  A sweep evaluates the three
// constraints in parallel; the values one lane needs from another (the
// two atoms of its constraint, the corrections of the two constraints an
// atom belongs to, the largest residual) move with DPP quad permutations,
// and every sum keeps the reference's order (constraint 0, 1, 2), so the
// results are those of the sequential loop bit for bit.
#include "mdg_shake.h"

namespace {

// quad_perm selectors: lane j of the quad reads lane sel_j
constexpr int QP_A = 0xD0;     // [0,0,1,3]: atom a of constraint k
constexpr int QP_B = 0xE9;     // [1,2,2,3]: atom b of constraint k
constexpr int QP_L0 = 0xC0;    // [0,0,0,3]
constexpr int QP_L1 = 0xD5;    // [1,1,1,3]
constexpr int QP_L2 = 0xEA;    // [2,2,2,3]
constexpr int QP_X1 = 0xB1;    // [1,0,3,2]
constexpr int QP_X2 = 0x4E;    // [2,3,0,1]

template <int CTRL>
__device__ inline double qperm(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The quad's largest value.
__device__ inline double quad_max(double v)
{
    v = fmax(v, qperm<QP_X1>(v));
    return fmax(v, qperm<QP_X2>(v));
}

// One lane's share of a molecule.
struct Q3 {
    int k;              // 0..2 live, 3 idle
    int at;             // atom k
    int cid;            // constraint k
    double x, y, z, vx, vy, vz, rm, nat;
    double rc2, src2, ddx, ddy, ddz;
    double rma, rmb;    // inverse masses of constraint k's atoms
};

__device__ inline void q_load(const ShakeArgs &A, int g, int k, Q3 &Q)
{
    Q.k = k;
    Q.at = -1;
    Q.cid = -1;
    Q.x = Q.y = Q.z = Q.vx = Q.vy = Q.vz = 0.;
    Q.rm = Q.nat = 0.;
    Q.rc2 = 1.;
    Q.src2 = 1.;
    Q.ddx = Q.ddy = Q.ddz = 0.;
    if (k < 3) {
        const int i = A.atoms[A.atomOff[g] + k];
        Q.at = i;
        Q.x = A.x[i];
        Q.y = A.y[i];
        Q.z = A.z[i];
        Q.vx = A.vx[i];
        Q.vy = A.vy[i];
        Q.vz = A.vz[i];
        Q.rm = A.rmass[i];
        Q.nat = 1.0 / (double)A.nAtConst[i];
        Q.cid = A.cons[A.conOff[g] + k];
        Q.rc2 = A.rc2[Q.cid];
        Q.src2 = sqrt(Q.rc2);
    }
    // loop invariants, moved once here (the moves cannot be hoisted by the
    // compiler out of the divergent sweep loops)
    Q.rma = qperm<QP_A>(Q.rm);
    Q.rmb = qperm<QP_B>(Q.rm);
}

__device__ inline void q_store(const ShakeArgs &A, const Q3 &Q, bool pos)
{
    if (Q.k >= 3) return;
    const int i = Q.at;
    if (pos) {
        A.x[i] = Q.x;
        A.y[i] = Q.y;
        A.z[i] = Q.z;
    }
    A.vx[i] = Q.vx;
    A.vy[i] = Q.vy;
    A.vz[i] = Q.vz;
}

__device__ inline void q_scr_store(const ShakeArgs &A, const Q3 &Q, int slot, bool pos)
{
    if (Q.k >= 3) return;
    const long n = A.P.n;
    double *s = A.scr + (long)slot * 6 * n;
    const int i = Q.at;
    if (pos) {
        s[i] = Q.x;
        s[n + i] = Q.y;
        s[2 * n + i] = Q.z;
    }
    s[3 * n + i] = Q.vx;
    s[4 * n + i] = Q.vy;
    s[5 * n + i] = Q.vz;
}

__device__ inline void q_scr_load(const ShakeArgs &A, Q3 &Q, int slot, bool pos)
{
    if (Q.k >= 3) return;
    const long n = A.P.n;
    const double *s = A.scr + (long)slot * 6 * n;
    const int i = Q.at;
    if (pos) {
        Q.x = s[i];
        Q.y = s[n + i];
        Q.z = s[2 * n + i];
    }
    Q.vx = s[3 * n + i];
    Q.vy = s[4 * n + i];
    Q.vz = s[5 * n + i];
}

// Minimum image of a short in-molecule vector: the reference's nint is 0
// whenever |d u| < 0.5, leaving d unchanged, so that case skips the rest.
__device__ inline double w3_dist(const MdgBox &B, double &dx, double &dy, double &dz)
{
#pragma clang fp contract(off)
    if (B.type == 1 && fabs(dx * B.u1) < 0.5 && fabs(dy * B.u1) < 0.5 && fabs(dz * B.u1) < 0.5)
        return dx * dx + dy * dy + dz * dz;
    return mdg_dist(B, dx, dy, dz);
}

// x_b - x_a of constraint k (every lane takes part in the moves).
__device__ inline void q_bond(const Q3 &Q, double &dx, double &dy, double &dz)
{
    dx = qperm<QP_B>(Q.x) - qperm<QP_A>(Q.x);
    dy = qperm<QP_B>(Q.y) - qperm<QP_A>(Q.y);
    dz = qperm<QP_B>(Q.z) - qperm<QP_A>(Q.z);
}

// Constraint vectors at the stage's start (integrate.c:1536-1545 with
// image_array), kept per constraint for the replay kernels.
__device__ inline void q_dd(const ShakeArgs &A, const MdgBox &B, Q3 &Q)
{
    double dx, dy, dz;
    q_bond(Q, dx, dy, dz);
    if (Q.k < 3) {
        w3_dist(B, dx, dy, dz);
        Q.ddx = dx;
        Q.ddy = dy;
        Q.ddz = dz;
        A.dd[Q.cid] = dx;
        A.dd[A.nConst + Q.cid] = dy;
        A.dd[2 * A.nConst + Q.cid] = dz;
    }
}

__device__ inline void q_dd_load(const ShakeArgs &A, Q3 &Q)
{
    if (Q.k >= 3) return;
    Q.ddx = A.dd[Q.cid];
    Q.ddy = A.dd[A.nConst + Q.cid];
    Q.ddz = A.dd[2 * A.nConst + Q.cid];
}

// shake.c:243-258: this lane's constraint vector and |r|^2; returns the
// molecule's 0.5 * max |r^2 - d^2| / d.
__device__ inline double q_check(const MdgBox &B, const Q3 &Q, double &dtx, double &dty,
                                 double &dtz, double &rt2)
{
#pragma clang fp contract(off)
    q_bond(Q, dtx, dty, dtz);
    double d = 0.;
    rt2 = 0.;
    if (Q.k < 3) {
        rt2 = w3_dist(B, dtx, dty, dtz);
        d = fabs(rt2 - Q.rc2) / Q.src2;
    }
    return 0.5 * quad_max(d);
}

// Each atom's two corrections in constraint order: atom 0 gets (c0, a)
// then (c1, a); atom 1 (c0, b) then (c2, a); atom 2 (c1, b) then (c2, b);
// li / lj are this lane's constraint's a-side / b-side terms.
__device__ inline double q_gather(int k, double li, double lj)
{
#pragma clang fp contract(off)
    const double a0 = qperm<QP_L0>(li), b0 = qperm<QP_A>(lj);
    const double a1 = qperm<QP_B>(li), b1 = qperm<QP_L2>(lj);
    const double first = k == 0 ? a0 : b0;
    const double second = k == 2 ? b1 : a1;
    double t = 0.;
    t += first;
    t += second;
    return t;
}

// v + t0 + t1 + t2 over the three live lanes, in that order, on every lane.
__device__ inline double q_sum3(double v, double t)
{
#pragma clang fp contract(off)
    const double t0 = qperm<QP_L0>(t), t1 = qperm<QP_L1>(t), t2 = qperm<QP_L2>(t);
    return ((v + t0) + t1) + t2;
}

// shake.c:262-332: one position sweep (lambda per constraint, corrections
// summed per atom, applied once per constraint membership).
__device__ inline void q_apply_r(const ShakeArgs &A, Q3 &Q, double dtx, double dty, double dtz,
                                 double rt2, double &vir)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double rma = Q.rma, rmb = Q.rmb;
    double lix = 0., liy = 0., liz = 0., ljx = 0., ljy = 0., ljz = 0., tv = 0.;
    if (Q.k < 3) {
        const double trmi = dt * rma, trmj = dt * rmb;
        const double lambda = -(Q.rc2 - rt2) /
            (dt * (trmi + trmj) * ((Q.ddx * dtx) + (Q.ddy * dty) + (Q.ddz * dtz)));
        tv = lambda * (Q.ddx * Q.ddx + Q.ddy * Q.ddy + Q.ddz * Q.ddz);
        const double li = 0.5 * lambda * trmi, lj = -0.5 * lambda * trmj;
        lix = Q.ddx * li;
        liy = Q.ddy * li;
        liz = Q.ddz * li;
        ljx = Q.ddx * lj;
        ljy = Q.ddy * lj;
        ljz = Q.ddz * lj;
    }
    vir = q_sum3(vir, tv);
    const double xt = q_gather(Q.k, lix, ljx), yt = q_gather(Q.k, liy, ljy),
                 zt = q_gather(Q.k, liz, ljz);
    if (Q.k < 3) {
        for (int m = 0; m < 2; m++) {   // the atom's two constraint memberships
            Q.x += dt * xt * Q.nat;
            Q.y += dt * yt * Q.nat;
            Q.z += dt * zt * Q.nat;
        }
        for (int m = 0; m < 2; m++) {
            Q.vx += xt * Q.nat;
            Q.vy += yt * Q.nat;
            Q.vz += zt * Q.nat;
        }
    }
}

// (trmi + trmj) |dd|^2 of this lane's constraint (shake.c:391-393).
__device__ inline double q_den(const ShakeArgs &A, const Q3 &Q)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double rma = Q.rma, rmb = Q.rmb;
    const double trmi = 0.5 * dt * rma, trmj = 0.5 * dt * rmb;
    return (trmi + trmj) * (Q.ddx * Q.ddx + Q.ddy * Q.ddy + Q.ddz * Q.ddz);
}

// shake.c:378-413: velocity lambdas of one sweep; returns the molecule's
// max |lambda|, the per-atom corrections in xt, yt, zt.
__device__ inline double q_lambda_v(const ShakeArgs &A, const Q3 &Q, double den, double &xt,
                                    double &yt, double &zt)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double rma = Q.rma, rmb = Q.rmb;
    const double dvx = qperm<QP_B>(Q.vx) - qperm<QP_A>(Q.vx);
    const double dvy = qperm<QP_B>(Q.vy) - qperm<QP_A>(Q.vy);
    const double dvz = qperm<QP_B>(Q.vz) - qperm<QP_A>(Q.vz);
    double lix = 0., liy = 0., liz = 0., ljx = 0., ljy = 0., ljz = 0., al = 0.;
    if (Q.k < 3) {
        const double trmi = 0.5 * dt * rma, trmj = 0.5 * dt * rmb;
        const double lambda = (Q.ddx * dvx + Q.ddy * dvy + Q.ddz * dvz) / den;
        al = fabs(lambda);
        const double li = lambda * trmi, lj = -lambda * trmj;
        lix = Q.ddx * li;
        liy = Q.ddy * li;
        liz = Q.ddz * li;
        ljx = Q.ddx * lj;
        ljy = Q.ddy * lj;
        ljz = Q.ddz * lj;
    }
    xt = q_gather(Q.k, lix, ljx);
    yt = q_gather(Q.k, liy, ljy);
    zt = q_gather(Q.k, liz, ljz);
    return quad_max(al);
}

// shake.c:418-436: one sweep's corrections, once per constraint membership.
__device__ inline void q_apply_v(Q3 &Q, double xt, double yt, double zt)
{
#pragma clang fp contract(off)
    if (Q.k >= 3) return;
    for (int m = 0; m < 2; m++) {
        Q.vx += xt * Q.nat;
        Q.vy += yt * Q.nat;
        Q.vz += zt * Q.nat;
    }
}

__device__ inline void q_kick(const ShakeArgs &A, Q3 &Q)
{
#pragma clang fp contract(off)
    if (Q.k >= 3) return;
    const double dt = A.P.timeStep;
    const int i = Q.at;
    Q.vx += 0.5 * dt * A.fx[i] * Q.rm;
    Q.vy += 0.5 * dt * A.fy[i] * Q.rm;
    Q.vz += 0.5 * dt * A.fz[i] * Q.rm;
}

// The box of molecule g: the run's, or its replica's in a batch.
template <bool BATCH>
__device__ inline const MdgBox &q_box(const ShakeArgs &A, int g)
{
    return BATCH ? A.boxes[A.grep[g]] : A.B;
}

template <bool BATCH>
__device__ inline const unsigned long long *q_mask(const ShakeArgs &A, int g)
{
    return BATCH ? A.mask + (long)A.grep[g] * SHK_MASKW : A.mask;
}

template <bool BATCH>
__device__ inline void q_publish(const ShakeArgs &A, int g, const unsigned long long (&bits)[SHK_MASKW],
                                 int kend, bool settled)
{
    if (BATCH) shake_publish_mask_b(A, bits, kend, settled, g < A.nGrp ? A.grep[g] : -1);
    else shake_publish_mask(A, bits, kend, settled);
}

}  // namespace

// Stage 1, part A: half kick, drift, SHAKE-R sweeps of each molecule on
// its own, recording the sweeps at which it was converged and the states
// of a window after its first convergence.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv1_a(ShakeArgs A)
{
#pragma clang fp contract(off)
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    // every lane reaches the wave-wide AND; idle molecules contribute all ones
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = q_box<BATCH>(A, g);
        Q3 Q;
        q_load(A, g, k, Q);
        const double dt = A.P.timeStep;
        q_dd(A, B, Q);
        q_kick(A, Q);
        if (k < 3) {
            Q.x += dt * Q.vx;
            Q.y += dt * Q.vy;
            Q.z += dt * Q.vz;
        }
        q_scr_store(A, Q, 0, true);
        for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
        double vir = 0.;
        int first = -1, it = 0;
        bool prev = false;
        settled = false;
        for (;;) {
            double dtx, dty, dtz, rt2;
            const bool cv = q_check(B, Q, dtx, dty, dtz, rt2) < A.P.tolShake;
            if (cv) bits[it >> 6] |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) {
                q_scr_store(A, Q, 1 + it - first, true);
                if (k == 0) A.virScr[(long)(1 + it - first) * A.nGrp + g] = vir;
            }
            if (cv && prev) {
                settled = true;
                break;
            }
            prev = cv;
            if (it + 1 >= A.P.maxCycle) break;
            q_apply_r(A, Q, dtx, dty, dtz, rt2, vir);
            it++;
        }
        q_scr_store(A, Q, SHK_W + 1, true);
        if (k == 0) {
            A.virScr[(long)(SHK_W + 1) * A.nGrp + g] = vir;
            A.first[g] = first;
            A.kend[g] = it;
        }
        kend = it;
    }
    q_publish<BATCH>(A, g, bits, kend, settled);
}

// Stage 1, part B: every molecule continues (or restarts from a saved
// state) to the first sweep C at which all molecules were converged.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv1_b(ShakeArgs A)
{
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    if (g >= A.nGrp) return;
    const int C = shake_global_stop(A, q_mask<BATCH>(A, g));
    if (C < 0) {
        if (k == 0) atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 1);
        return;
    }
    const MdgBox &B = q_box<BATCH>(A, g);
    const int first = A.first[g], kend = A.kend[g];
    int slot, at;
    shake_pick_slot(C, first, kend, slot, at);
    Q3 Q;
    q_load(A, g, k, Q);
    q_dd_load(A, Q);
    q_scr_load(A, Q, slot, true);
    double vir = slot == 0 ? 0. : A.virScr[(long)slot * A.nGrp + g];
    for (int it = at; it < C; it++) {
        double dtx, dty, dtz, rt2;
        q_check(B, Q, dtx, dty, dtz, rt2);
        q_apply_r(A, Q, dtx, dty, dtz, rt2, vir);
    }
    if (k == 0) A.vir[g] = vir;
    q_store(A, Q, true);
}

// Stage 2, part A: constraint vectors, half kick, SHAKE-V sweeps with the
// same bookkeeping as stage 1.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv2_a(ShakeArgs A)
{
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = q_box<BATCH>(A, g);
        Q3 Q;
        q_load(A, g, k, Q);
        q_dd(A, B, Q);
        q_kick(A, Q);
        q_scr_store(A, Q, 0, false);
        const double den = q_den(A, Q);
        const double tolvel = A.P.tolShake * A.P.rTimeStep;
        for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
        int first = -1, it = 0;
        bool prev = false;
        settled = false;
        for (;;) {
            // the check is the sweep's own max |lambda| (shake.c:421); every
            // sweep that is not the last one checked is applied, because the
            // reference's global loop goes on while any group is unconverged
            double xt, yt, zt;
            const bool cv = q_lambda_v(A, Q, den, xt, yt, zt) < tolvel;
            if (cv) bits[it >> 6] |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) q_scr_store(A, Q, 1 + it - first, false);
            if (cv && prev) {
                settled = true;
                break;
            }
            prev = cv;
            if (it + 1 >= A.P.maxCycle) break;
            q_apply_v(Q, xt, yt, zt);
            it++;
        }
        q_scr_store(A, Q, SHK_W + 1, false);
        if (k == 0) {
            A.first[g] = first;
            A.kend[g] = it;
        }
        kend = it;
    }
    q_publish<BATCH>(A, g, bits, kend, settled);
}

// Stage 2, part B: replay to the global sweep C, kinetic energy and stress
// partials (utils.c:502-515), wrap (utils.c:162-218).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv2_b(ShakeArgs A)
{
#pragma clang fp contract(off)
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (g < A.nGrp) {
        const MdgBox &B = q_box<BATCH>(A, g);
        Q3 Q;
        q_load(A, g, k, Q);
        const int C = shake_global_stop(A, q_mask<BATCH>(A, g));
        if (C < 0) {
            if (k == 0) atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 2);
        } else {
            const int first = A.first[g], kend = A.kend[g];
            int slot, at;
            shake_pick_slot(C, first, kend, slot, at);
            q_dd_load(A, Q);
            q_scr_load(A, Q, slot, false);
            const double den = q_den(A, Q);
            for (int it = at; it < C; it++) {
                double xt, yt, zt;
                q_lambda_v(A, Q, den, xt, yt, zt);
                q_apply_v(Q, xt, yt, zt);
            }
        }
        double ke = 0.;
        if (k < 3) {
            const double m = A.mass[Q.at];
            ke = m * (Q.vx * Q.vx + Q.vy * Q.vy + Q.vz * Q.vz);
            acc[0] = ke;
            acc[1] = m * Q.vx * Q.vx;
            acc[2] = m * Q.vx * Q.vy;
            acc[3] = m * Q.vx * Q.vz;
            acc[4] = m * Q.vy * Q.vy;
            acc[5] = m * Q.vy * Q.vz;
            acc[6] = m * Q.vz * Q.vz;
            shake_image(B, Q.x, Q.y, Q.z);
        }
        const double kin = q_sum3(0., ke);   // the molecule's sum, in atom order
        q_store(A, Q, true);
        if (BATCH && k == 0) A.grpKin[g] = kin;
    }
    mdg_block_partials<7>(acc, A.part);
}

template __global__ void k_w3_vv1_a<false>(ShakeArgs);
template __global__ void k_w3_vv1_a<true>(ShakeArgs);
template __global__ void k_w3_vv1_b<false>(ShakeArgs);
template __global__ void k_w3_vv1_b<true>(ShakeArgs);
template __global__ void k_w3_vv2_a<false>(ShakeArgs);
template __global__ void k_w3_vv2_a<true>(ShakeArgs);
template __global__ void k_w3_vv2_b<false>(ShakeArgs);
template __global__ void k_w3_vv2_b<true>(ShakeArgs);
