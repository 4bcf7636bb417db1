// mdg_shake_w3.hip -- VV + SHAKE kernels specialised for rigid
// three-site molecules (three atoms a0 < a1 < a2 with constraints
// (a0,a1), (a0,a2), (a1,a2) in that order, e.g. TIP3P water under
// consth).  Same arithmetic, operation order and global stopping rule as
// the generic kernels of mdg_integrate.hip.
//
// One molecule per quad of lanes: lane k < 3 holds atom k's state and the
// data of one constraint that contains atom k, lane 3 idles.  Molecules sit on every spread-th
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
constexpr int QP_R = 0xC9;     // [1,2,0,3]: the other atom of the lane's constraint
constexpr int QP_A = 0xC4;     // [0,1,0,3]: atom a of the lane's constraint
constexpr int QP_B = 0xE9;     // [1,2,2,3]: atom b of the lane's constraint
constexpr int QP_CA = 0xE0;    // [0,0,2,3]: lane of atom k's first constraint
constexpr int QP_CB = 0xD6;    // [2,1,1,3]: lane of atom k's second constraint
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

template <int CTRL>
__device__ inline int qperm_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

// True on every lane of the quad when c holds on all four lanes.
__device__ inline bool quad_all(bool c)
{
    int v = c ? 1 : 0;
    v &= qperm_i<QP_X1>(v);
    v &= qperm_i<QP_X2>(v);
    return v != 0;
}

// One lane's share of a molecule.  Lane k < 3 holds atom k and one
// constraint that contains atom k: lane 0 constraint 0 = (0,1), lane 1
// constraint 2 = (1,2), lane 2 constraint 1 = (0,2); lane 3 idles with zero
// data.  A lane forms its constraint's vector as (the other atom) - (its
// own atom) with one DPP read per component: x_b - x_a on lanes 0 and 1,
// and -(x_b - x_a), exactly, on lane 2.  That sign cancels wherever the
// vector enters: |d|^2, the image shift (nint is odd), the multiplier
// (dd . dt with both vectors negated), and the corrections, whose factors
// for constraint 1 are negated below, so every product is the reference's.
//
// Correction exchange: atom 0 receives the a-side terms of constraints 0
// and 1, atom 1 the b-side of 0 and the a-side of 2, atom 2 the b-side of 1
// and 2: in constraint order, its first constraint cA sits on lane
// (0, 0, 2)[k] and its second cB on lane (2, 1, 1)[k] (QP_CA, QP_CB).
// Each sweep only the constraints' multipliers move (two DPP reads); the
// atom holds the vectors dd of cA and cB and its factors fA, fB (trmi, or
// -trmj on the b side; negated for constraint 1) from the stage's start,
// forms li = h fA and lj = h fB as the constraint's lane would, and sums
// (0 + dd_cA l_A) + dd_cB l_B as the reference accumulates xt[]
// (shake.c:300-310).
struct Q3 {
    int k;              // 0..2 live, 3 idle
    int at;             // atom k
    int cid;            // the lane's constraint (0, 2, 1)[k]
    double x, y, z, vx, vy, vz, rm, nat;
    double rc2, thr;    // stage 1: converged while |r^2 - rc2| < thr
    double ddx, ddy, ddz, ddsq;   // constraint vector at the stage start, |dd|^2
    double shx, shy, shz;         // image shift of the constraint vector (stage 1)
    double K;           // stage 1: dt (trmi + trmj); stage 2: (trmi + trmj) |dd|^2
    double fA, fB;      // the atom's factors in its constraints cA, cB
    double dAx, dAy, dAz, dBx, dBy, dBz;   // dd of cA and cB
    double rma, rmb;    // inverse masses of the lane constraint's atoms a, b
};

__device__ inline void q_load(const ShakeArgs &A, int g, int k, Q3 &Q)
{
    Q.k = k;
    Q.at = -1;
    Q.cid = -1;
    Q.x = Q.y = Q.z = Q.vx = Q.vy = Q.vz = 0.;
    Q.rm = Q.nat = 0.;
    Q.rc2 = 0.;
    Q.thr = 1.;
    Q.ddx = Q.ddy = Q.ddz = Q.ddsq = 0.;
    Q.shx = Q.shy = Q.shz = 0.;
    Q.K = 1.;
    Q.fA = Q.fB = 0.;
    Q.dAx = Q.dAy = Q.dAz = Q.dBx = Q.dBy = Q.dBz = 0.;
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
        Q.cid = A.cons[A.conOff[g] + (k == 0 ? 0 : 3 - k)];
        Q.rc2 = A.rc2[Q.cid];
        Q.thr = A.thr[Q.cid];
    }
    Q.rma = qperm<QP_A>(Q.rm);
    Q.rmb = qperm<QP_B>(Q.rm);
}

// The atom's factors and constraint vectors from its constraint lanes'
// a-side and b-side coefficients ti, tj (li = h ti, lj = -h tj for the
// sweep's multiplier h): atom 0 is the a side of constraints 0 and 1, atom
// 1 the b side of 0 and the a side of 2, atom 2 the b side of 1 and 2.
__device__ inline void q_factors(Q3 &Q, double ti, double tj)
{
    const double aI = qperm<QP_CA>(ti), aJ = qperm<QP_CA>(tj);
    const double bI = qperm<QP_CB>(ti), bJ = qperm<QP_CB>(tj);
    // sides as above; constraint 1 (atom 0's cB, atom 2's cA) is held
    // negated, so its factors are too
    Q.fA = Q.k == 0 ? aI : Q.k == 1 ? -aJ : aJ;
    Q.fB = Q.k == 0 ? -bI : Q.k == 1 ? bI : -bJ;
    Q.dAx = qperm<QP_CA>(Q.ddx);
    Q.dAy = qperm<QP_CA>(Q.ddy);
    Q.dAz = qperm<QP_CA>(Q.ddz);
    Q.dBx = qperm<QP_CB>(Q.ddx);
    Q.dBy = qperm<QP_CB>(Q.ddy);
    Q.dBz = qperm<QP_CB>(Q.ddz);
    if (Q.k >= 3) Q.fA = Q.fB = 0.;
}

// Stage 1: lambda denominators and factors (trmi = dt rma, trmj = dt rmb;
// li = 0.5 lambda trmi, lj = -0.5 lambda trmj, shake.c:262-332).
__device__ inline void q_setup_r(const ShakeArgs &A, Q3 &Q)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double trmi = dt * Q.rma, trmj = dt * Q.rmb;
    Q.K = Q.k < 3 ? dt * (trmi + trmj) : 1.;
    Q.ddsq = Q.ddx * Q.ddx + Q.ddy * Q.ddy + Q.ddz * Q.ddz;
    q_factors(Q, trmi, trmj);
}

// Stage 2: den = (trmi + trmj) |dd|^2 with trmi = 0.5 dt rma, trmj = 0.5 dt
// rmb; li = lambda trmi, lj = -lambda trmj (shake.c:378-413).
__device__ inline void q_setup_v(const ShakeArgs &A, Q3 &Q)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double trmi = 0.5 * dt * Q.rma, trmj = 0.5 * dt * Q.rmb;
    Q.K = Q.k < 3 ? (trmi + trmj) * (Q.ddx * Q.ddx + Q.ddy * Q.ddy + Q.ddz * Q.ddz) : 1.;
    q_factors(Q, trmi, trmj);
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
    double *s = A.scr + (long)slot * 6 * n + Q.at;
    if (pos) {
        s[0] = Q.x;
        s[n] = Q.y;
        s[2 * n] = Q.z;
    }
    s[3 * n] = Q.vx;
    s[4 * n] = Q.vy;
    s[5 * n] = Q.vz;
}

__device__ inline void q_scr_load(const ShakeArgs &A, Q3 &Q, int slot, bool pos)
{
    if (Q.k >= 3) return;
    const long n = A.P.n;
    const double *s = A.scr + (long)slot * 6 * n + Q.at;
    if (pos) {
        Q.x = s[0];
        Q.y = s[n];
        Q.z = s[2 * n];
    }
    Q.vx = s[3 * n];
    Q.vy = s[4 * n];
    Q.vz = s[5 * n];
}

// The lane constraint's vector, x_b - x_a (lane 2: its negation); every
// lane takes part in the move.
__device__ inline void q_bond(const Q3 &Q, double &dx, double &dy, double &dz)
{
    dx = qperm<QP_R>(Q.x) - Q.x;
    dy = qperm<QP_R>(Q.y) - Q.y;
    dz = qperm<QP_R>(Q.z) - Q.z;
}

// The image shift utils.c:38-80 subtracts from a constraint vector, as a
// constant vector: a bond (about an angstrom) keeps its shift over the
// sweeps of a stage, and d - shift is then the reference's expression.
__device__ inline void q_shift(const MdgBox &B, Q3 &Q, double dx, double dy, double dz)
{
#pragma clang fp contract(off)
    if (Q.k >= 3 || B.type == 3) return;
    const double Ly = B.type == 1 ? B.a1 : B.b2, Lz = B.type == 1 ? B.a1 : B.c3;
    const double uy = B.type == 1 ? B.u1 : B.v2, uz = B.type == 1 ? B.u1 : B.w3;
    Q.shx = B.a1 * mdg_nint(dx * B.u1);
    Q.shy = Ly * mdg_nint(dy * uy);
    Q.shz = Lz * mdg_nint(dz * uz);
}

// Constraint vector with the stage's image shift and its |d|^2 as
// (dx^2 + dy^2) + dz^2 (triclinic boxes take the full image every sweep).
__device__ inline double q_img(const MdgBox &B, const Q3 &Q, double &dx, double &dy, double &dz)
{
#pragma clang fp contract(off)
    if (B.type == 3) return mdg_dist(B, dx, dy, dz);
    dx = dx - Q.shx;
    dy = dy - Q.shy;
    dz = dz - Q.shz;
    return dx * dx + dy * dy + dz * dz;
}

// Constraint vectors at the stage's start (integrate.c:1536-1545 with
// image_array), kept per constraint for the replay kernels.
__device__ inline void q_dd(const ShakeArgs &A, const MdgBox &B, Q3 &Q)
{
    double dx, dy, dz;
    q_bond(Q, dx, dy, dz);
    if (Q.k < 3) {
        mdg_dist(B, dx, dy, dz);
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

// The image shift of this lane's constraint at the current positions.
__device__ inline void q_shift_now(const MdgBox &B, Q3 &Q)
{
    double dx, dy, dz;
    q_bond(Q, dx, dy, dz);
    q_shift(B, Q, dx, dy, dz);
}

// shake.c:243-258: the molecule is converged when every constraint has
// |r^2 - d^2| below its threshold; dt*, rt2 keep this lane's vector.
__device__ inline bool q_check(const MdgBox &B, const Q3 &Q, double &dtx, double &dty,
                               double &dtz, double &rt2)
{
#pragma clang fp contract(off)
    q_bond(Q, dtx, dty, dtz);
    rt2 = q_img(B, Q, dtx, dty, dtz);
    return quad_all(Q.k >= 3 || fabs(rt2 - Q.rc2) < Q.thr);
}

// The atom's corrections from its two constraints' multipliers h (every
// lane takes part in the moves).
__device__ inline void q_exchange(const Q3 &Q, double h, double &xt, double &yt, double &zt)
{
#pragma clang fp contract(off)
    const double lA = qperm<QP_CA>(h) * Q.fA, lB = qperm<QP_CB>(h) * Q.fB;
    xt = (0. + Q.dAx * lA) + Q.dBx * lB;
    yt = (0. + Q.dAy * lA) + Q.dBy * lB;
    zt = (0. + Q.dAz * lA) + Q.dBz * lB;
}

// v + t0 + t1 + t2 over the three live lanes, in that order, on every lane.
__device__ inline double q_sum3(double v, double t)
{
#pragma clang fp contract(off)
    const double t0 = qperm<QP_L0>(t), t1 = qperm<QP_L1>(t), t2 = qperm<QP_L2>(t);
    return ((v + t0) + t1) + t2;
}

// The same over the constraints in their order 0, 1, 2 (lanes 0, 2, 1).
__device__ inline double q_sum3c(double v, double t)
{
#pragma clang fp contract(off)
    const double t0 = qperm<QP_L0>(t), t1 = qperm<QP_L2>(t), t2 = qperm<QP_L1>(t);
    return ((v + t0) + t1) + t2;
}

// The state after one sweep, formed before the sweep's stopping decision
// and taken only when the sweep is applied: the multiplier's division and
// the corrections then overlap the wave-wide vote instead of following it.
struct QNext {
    double lambda;
    double x, y, z, vx, vy, vz;
};

// Keeps the values computed at this point of the sweep: an empty asm
// statement that reads them, which the compiler neither moves nor deletes,
// so it cannot sink their arithmetic past the branch that follows.
__device__ inline void q_pin(const QNext &N)
{
    asm volatile("" ::"v"(N.x), "v"(N.y), "v"(N.z), "v"(N.vx), "v"(N.vy), "v"(N.vz));
}

// shake.c:262-332: one position sweep (lambda per constraint, corrections
// summed per atom, applied once per constraint membership).
__device__ inline void q_next_r(const ShakeArgs &A, const Q3 &Q, double dtx, double dty,
                                double dtz, double rt2, QNext &N)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    const double dot = Q.ddx * dtx + Q.ddy * dty + Q.ddz * dtz;
    double lambda = -(Q.rc2 - rt2) / (Q.K * dot);
    lambda = Q.k < 3 ? lambda : 0.;
    N.lambda = lambda;
    double xt, yt, zt;
    q_exchange(Q, 0.5 * lambda, xt, yt, zt);
    // the correction once per constraint membership (two per atom); the
    // idle lane has nat = 0 and zero corrections
    const double ix = dt * xt * Q.nat, iy = dt * yt * Q.nat, iz = dt * zt * Q.nat;
    N.x = (Q.x + ix) + ix;
    N.y = (Q.y + iy) + iy;
    N.z = (Q.z + iz) + iz;
    const double jx = xt * Q.nat, jy = yt * Q.nat, jz = zt * Q.nat;
    N.vx = (Q.vx + jx) + jx;
    N.vy = (Q.vy + jy) + jy;
    N.vz = (Q.vz + jz) + jz;
}

// Applies a position sweep; vir collects this lane's constraint's virial
// terms.
__device__ inline void q_take_r(const QNext &N, Q3 &Q, double &vir)
{
#pragma clang fp contract(off)
    vir += N.lambda * Q.ddsq;
    Q.x = N.x;
    Q.y = N.y;
    Q.z = N.z;
    Q.vx = N.vx;
    Q.vy = N.vy;
    Q.vz = N.vz;
}

__device__ inline void q_apply_r(const ShakeArgs &A, Q3 &Q, double dtx, double dty, double dtz,
                                 double rt2, double &vir)
{
    QNext N;
    q_next_r(A, Q, dtx, dty, dtz, rt2, N);
    q_take_r(N, Q, vir);
}

// shake.c:378-413: one velocity sweep's lambda; converged when every
// |lambda| of the molecule is below tolvel.  Corrections in xt, yt, zt.
__device__ inline bool q_lambda_v(const Q3 &Q, double tolvel, double &xt, double &yt, double &zt)
{
#pragma clang fp contract(off)
    const double dvx = qperm<QP_R>(Q.vx) - Q.vx;
    const double dvy = qperm<QP_R>(Q.vy) - Q.vy;
    const double dvz = qperm<QP_R>(Q.vz) - Q.vz;
    double lambda = (Q.ddx * dvx + Q.ddy * dvy + Q.ddz * dvz) / Q.K;
    lambda = Q.k < 3 ? lambda : 0.;
    q_exchange(Q, lambda, xt, yt, zt);
    return quad_all(fabs(lambda) < tolvel);
}

// shake.c:418-436: one sweep's corrections, once per constraint membership
// (positions are not moved in this stage: N keeps them).
__device__ inline void q_next_v(const Q3 &Q, double xt, double yt, double zt, QNext &N)
{
#pragma clang fp contract(off)
    const double jx = xt * Q.nat, jy = yt * Q.nat, jz = zt * Q.nat;
    N.lambda = 0.;
    N.x = Q.x;
    N.y = Q.y;
    N.z = Q.z;
    N.vx = (Q.vx + jx) + jx;
    N.vy = (Q.vy + jy) + jy;
    N.vz = (Q.vz + jz) + jz;
}

__device__ inline void q_apply_v(Q3 &Q, double xt, double yt, double zt)
{
    QNext N;
    q_next_v(Q, xt, yt, zt, N);
    Q.vx = N.vx;
    Q.vy = N.vy;
    Q.vz = N.vz;
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

// Converged-sweep bits: the current 64-sweep word in a register, flushed to
// the array when the word is complete or the loop ends.
struct QBits {
    unsigned long long w[SHK_MASKW];
    unsigned long long cur;
};

__device__ inline void qb_flush(QBits &b, int word)
{
    for (int q = 0; q < SHK_MASKW; q++)
        if (q == word) b.w[q] = b.cur;
    b.cur = 0;
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
// of a window after its first convergence.  Blocks past S.nbMain do the
// previous step's record and list prune (mdg_side.h).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv1_a(ShakeArgs A, SideWork S)
{
#pragma clang fp contract(off)
    if ((int)blockIdx.x >= S.nbMain) {
        side_work<BATCH>(S);
        return;
    }
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    // every lane reaches the wave-wide AND; idle molecules contribute all ones
    QBits bits;
    for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = ~0ull;
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = q_box<BATCH>(A, g);
        Q3 Q;
        q_load(A, g, k, Q);
        const double dt = A.P.timeStep;
        q_dd(A, B, Q);
        q_setup_r(A, Q);
        q_kick(A, Q);
        if (k < 3) {
            Q.x += dt * Q.vx;
            Q.y += dt * Q.vy;
            Q.z += dt * Q.vz;
        }
        q_shift_now(B, Q);
        q_scr_store(A, Q, 0, true);
        for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = 0;
        bits.cur = 0;
        double vir = 0.;
        int first = -1, it = 0, nr = 0;
        int rr[SHK_R] = {-1, -1, -1, -1};
        // one system: the wavefront sweeps as a whole; a batch (measured
        // slower that way), molecule by molecule
        const bool whole = !BATCH;
        bool prev = false;
        settled = false;
        // Molecules sweep on until every molecule of the wavefront has
        // settled (the wavefront's time is the same; the later sweeps are
        // the ones the replay would make).  A sweep at which the whole
        // wavefront is converged may be the global stop: its state goes to
        // the next of SHK_R ring slots.  In a replica batch each molecule
        // stops on its own (a wavefront can hold two replicas with different
        // stops, and the batch measured slower with whole-wavefront sweeps).
        for (;;) {
            double dtx, dty, dtz, rt2;
            const bool cv = q_check(B, Q, dtx, dty, dtz, rt2);
            QNext N;
            q_next_r(A, Q, dtx, dty, dtz, rt2, N);
            q_pin(N);
            if (cv) bits.cur |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) {
                q_scr_store(A, Q, 1 + it - first, true);
                if (k < 3) A.virScr[((long)(1 + it - first) * A.nGrp + g) * 3 + k] = vir;
            }
            if (whole && __builtin_amdgcn_ballot_w64(!cv) == 0) {
                q_scr_store(A, Q, SHK_W + 2 + (nr & (SHK_R - 1)), true);
                if (k < 3) A.virScr[((long)(SHK_W + 2 + (nr & (SHK_R - 1))) * A.nGrp + g) * 3 + k] = vir;
#pragma unroll
                for (int q = 0; q < SHK_R; q++)
                    if (q == (nr & (SHK_R - 1))) rr[q] = it;
                nr++;
            }
            if (whole ? __builtin_amdgcn_ballot_w64(!(cv && prev)) == 0 : cv && prev) {
                settled = true;
                break;
            }
            prev = cv;
            if (it + 1 >= A.P.maxCycle) break;
            q_take_r(N, Q, vir);
            {
                // inside the loop every live lane of the wavefront is at the
                // same sweep, so the word is uniform
                const int its = __builtin_amdgcn_readfirstlane(it);
                if ((its & 63) == 63) qb_flush(bits, its >> 6);
            }
            it++;
        }
        qb_flush(bits, it >> 6);
        q_scr_store(A, Q, SHK_W + 1, true);
        if (k < 3) A.virScr[((long)(SHK_W + 1) * A.nGrp + g) * 3 + k] = vir;
        if (k == 0) {
            A.first[g] = first;
            A.kend[g] = it;
            for (int q = 0; q < SHK_R; q++) A.ring[(long)SHK_R * g + q] = rr[q];
        }
        kend = it;
    }
    q_publish<BATCH>(A, g, bits.w, kend, settled);
}

// Stage 1, part B: every molecule continues (or restarts from a saved
// state) to the first sweep C at which all molecules were converged.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv1_b(ShakeArgs A)
{
    // the stage-2 mask (beside this stage's) for this step's stage 2
    if (!BATCH && blockIdx.x == 0 && threadIdx.x < SHK_MASKW) A.mask[SHK_MASKW + threadIdx.x] = ~0ull;
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    double vblk = 0.;   // this lane's share of the block's SHAKE virial
    int myAtom = -1;    // this lane's atom (lanes 0..2 of a live molecule)
    if (g < A.nGrp) {
        if (k < 3) myAtom = A.atoms[A.atomOff[g] + k];
        const int C = shake_global_stop(A, q_mask<BATCH>(A, g));
        if (C < 0) {
            if (k == 0) atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 1);
        } else {
            const MdgBox &B = q_box<BATCH>(A, g);
            const int first = A.first[g], kend = A.kend[g];
            int slot, at;
            shake_pick_slot(C, first, kend, A.ring + (long)SHK_R * g, slot, at);
            Q3 Q;
            q_load(A, g, k, Q);
            q_dd_load(A, Q);
            q_setup_r(A, Q);
            q_scr_load(A, Q, slot, true);
            q_shift_now(B, Q);
            // the saved virial is the molecule's; it continues on lane 0
            // each constraint continues its own running sum, so the result
            // does not depend on the slot the replay starts from
            double vir = (slot == 0 || k >= 3) ? 0. : A.virScr[((long)slot * A.nGrp + g) * 3 + k];
            for (int it = at; it < C; it++) {
                double dtx, dty, dtz, rt2;
                q_check(B, Q, dtx, dty, dtz, rt2);
                q_apply_r(A, Q, dtx, dty, dtz, rt2, vir);
            }
            {
                // the state at C passes the check, unless a molecule
                // re-diverged after it was taken as settled
                double dtx, dty, dtz, rt2;
                if (!q_check(B, Q, dtx, dty, dtz, rt2) && k == 0)
                    atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
            }
            const double vm = q_sum3c(0., vir);
            if (BATCH) {
                if (k == 0) A.vir[g] = vm;   // per molecule, summed per replica
            } else if (k == 0) {
                vblk = vm;
            }
            q_store(A, Q, true);
        }
    }
    // slot positions (the atom's own store above is visible to this lane)
    double d2 = 0., rel = 0.;
    const bool track = A.track && A.ctrl[CTRL_PRUNE] == 1;
    if (!BATCH && A.slotWrite && myAtom >= 0) {
        const int s = A.atomSlot[myAtom];
        const double nx = A.x[myAtom], ny = A.y[myAtom], nz = A.z[myAtom];
        if (A.spr) rel = mdg_slot_rel(A.B, A.cO, A.ncluO, s / MDG_CL, nx, ny, nz, A.spr, s);
        if (track) {
            const double L0 = A.B.a1, L1 = A.B.type == 1 ? A.B.a1 : A.B.b2,
                         L2 = A.B.type == 1 ? A.B.a1 : A.B.c3;
            double dx = nx - A.sx[s], dy = ny - A.sy[s], dz = nz - A.sz[s];
            dx -= L0 * rint(dx / L0);
            dy -= L1 * rint(dy / L1);
            dz -= L2 * rint(dz / L2);
            d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 <= 1e300)) d2 = INFINITY;   // NaN positions: no pruning
        }
        A.sx[s] = nx;
        A.sy[s] = ny;
        A.sz[s] = nz;
        A.fix[s] = 0;
        A.fix[s + A.fstr] = 0;
        A.fix[s + 2 * A.fstr] = 0;
    }
    if (!BATCH) {
        // one partial per block, summed in a fixed order by the record
        double acc[1] = {vblk};
        mdg_block_partials<1>(acc, A.vir);
    }
    if (!BATCH && A.slotWrite) {
        for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
        __shared__ double wmax[4];
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d2;
        __syncthreads();
        if (threadIdx.x == 0)
            A.dispPart[blockIdx.x] =
                track ? fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3])) : INFINITY;
        if (A.spr) mdg_block_max(rel, A.relPart);
    }
}

// Stage 2, part A: constraint vectors, half kick, SHAKE-V sweeps with the
// same bookkeeping as stage 1.
//
// With fuse set (1 mixed, 2 FP64; one box), the kernel first assembles each
// atom's force as k_finalize does (fin_atom, mdg_side.h: the same
// arithmetic, so k_finalize is not launched), keeps the correction
// partials per block, and resets the stage-1 SHAKE mask for the next step
// (k_w3_vv1_b resets the stage-2 one).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_w3_vv2_a(ShakeArgs A, FinArgs F, int fuse)
{
    __shared__ double mean[3], red[4][3];
    double facc[2] = {0., 0.};
    if (!BATCH && fuse) {
        fin_mean(F, mean, red);
        if (blockIdx.x == 0 && threadIdx.x < SHK_MASKW) F.mask[threadIdx.x] = ~0ull;
    }
    const int g = shake_w3_group(A);
    const int k = threadIdx.x & 3;
    QBits bits;
    for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = ~0ull;
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = q_box<BATCH>(A, g);
        Q3 Q;
        q_load(A, g, k, Q);
        if (!BATCH && fuse && k < 3) {
            double gx, gy, gz;
            const bool fr = fuse == 1 ? fin_atom<true>(F, Q.at, mean, facc, gx, gy, gz)
                                      : fin_atom<false>(F, Q.at, mean, facc, gx, gy, gz);
            if (fr) Q.vx = Q.vy = Q.vz = 0.;
        }
        q_dd(A, B, Q);
        q_setup_v(A, Q);
        q_kick(A, Q);
        q_scr_store(A, Q, 0, false);
        const double tolvel = A.P.tolShake * A.P.rTimeStep;
        for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = 0;
        bits.cur = 0;
        int first = -1, it = 0, nr = 0;
        int rr[SHK_R] = {-1, -1, -1, -1};
        // one system: the wavefront sweeps as a whole; a batch (measured
        // slower that way), molecule by molecule
        const bool whole = !BATCH;
        bool prev = false;
        settled = false;
        for (;;) {
            // the check is the sweep's own max |lambda| (shake.c:421); every
            // sweep that is not the last one checked is applied, because the
            // reference's global loop goes on while any group is unconverged.
            // The wavefront sweeps on until all its molecules have settled,
            // with the ring of wavefront-converged states, as in stage 1.
            double xt, yt, zt;
            const bool cv = q_lambda_v(Q, tolvel, xt, yt, zt);
            QNext N;
            q_next_v(Q, xt, yt, zt, N);
            q_pin(N);
            if (cv) bits.cur |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) q_scr_store(A, Q, 1 + it - first, false);
            if (whole && __builtin_amdgcn_ballot_w64(!cv) == 0) {
                q_scr_store(A, Q, SHK_W + 2 + (nr & (SHK_R - 1)), false);
#pragma unroll
                for (int q = 0; q < SHK_R; q++)
                    if (q == (nr & (SHK_R - 1))) rr[q] = it;
                nr++;
            }
            if (whole ? __builtin_amdgcn_ballot_w64(!(cv && prev)) == 0 : cv && prev) {
                settled = true;
                break;
            }
            prev = cv;
            if (it + 1 >= A.P.maxCycle) break;
            Q.vx = N.vx;
            Q.vy = N.vy;
            Q.vz = N.vz;
            {
                // inside the loop every live lane of the wavefront is at the
                // same sweep, so the word is uniform
                const int its = __builtin_amdgcn_readfirstlane(it);
                if ((its & 63) == 63) qb_flush(bits, its >> 6);
            }
            it++;
        }
        qb_flush(bits, it >> 6);
        q_scr_store(A, Q, SHK_W + 1, false);
        if (k == 0) {
            A.first[g] = first;
            A.kend[g] = it;
            for (int q = 0; q < SHK_R; q++) A.ring[(long)SHK_R * g + q] = rr[q];
        }
        kend = it;
    }
    q_publish<BATCH>(A, g, bits.w, kend, settled);
    if (!BATCH && fuse) mdg_block_partials<2>(facc, F.part);
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
            shake_pick_slot(C, first, kend, A.ring + (long)SHK_R * g, slot, at);
            q_dd_load(A, Q);
            q_setup_v(A, Q);
            q_scr_load(A, Q, slot, false);
            for (int it = at; it < C; it++) {
                double xt, yt, zt;
                q_lambda_v(Q, 0., xt, yt, zt);
                q_apply_v(Q, xt, yt, zt);
            }
            {
                double xt, yt, zt;
                if (!q_lambda_v(Q, A.P.tolShake * A.P.rTimeStep, xt, yt, zt) && k == 0)
                    atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
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

template __global__ void k_w3_vv1_a<false>(ShakeArgs, SideWork);
template __global__ void k_w3_vv1_a<true>(ShakeArgs, SideWork);
template __global__ void k_w3_vv1_b<false>(ShakeArgs);
template __global__ void k_w3_vv1_b<true>(ShakeArgs);
template __global__ void k_w3_vv2_a<false>(ShakeArgs, FinArgs, int);
template __global__ void k_w3_vv2_a<true>(ShakeArgs, FinArgs, int);
template __global__ void k_w3_vv2_b<false>(ShakeArgs);
template __global__ void k_w3_vv2_b<true>(ShakeArgs);
