// mdg_shake_m3.hip -- VV + SHAKE kernels for rigid three-site molecules
// with one molecule per lane (SHAKE spread 0).  Same molecules, stages,
// saved-state window and global stopping rule as mdg_shake_w3.hip, which
// holds one molecule per quad of lanes; here one lane holds the three atoms
// (a0 < a1 < a2) and the three constraints (a0,a1), (a0,a2), (a1,a2).
//
// A molecule's SHAKE sweeps are a chain of dependent FP64 operations, and
// the stopping rule fixes their number (tens per stage for water at
// tolShake 1e-8).  With a molecule per quad every sweep passes values
// between lanes (DPP moves) on the critical path; with a molecule per lane
// the three constraints are three independent chains in one lane, which
// the SIMD interleaves, and no values move between lanes.  A wavefront
// holds 64 molecules, so a box of 10,000 waters takes 157 wavefronts:
// fewer than the SIMDs, and each runs its sweeps at the issue rate rather
// than at the latency of the chain.
//
// Arithmetic is the reference's loop order (shake.c:262-332, 378-436):
// corrections summed per atom in constraint order from zero, applied once
// per constraint membership.  The results equal the quad kernels' bit for
// bit (their DPP sums are the same two-term sums).
#include "mdg_shake.h"

namespace {

// atoms a, b of constraint c
__device__ constexpr int M_A[3] = {0, 0, 1};
__device__ constexpr int M_B[3] = {1, 2, 2};

struct M3 {
    int at[3];          // atoms
    int cid[3];         // constraints
    double x[3], y[3], z[3], vx[3], vy[3], vz[3], rm[3], nat[3];
    double rc2[3], thr[3];
    double ddx[3], ddy[3], ddz[3], ddsq[3];   // constraint vectors at the stage start
    double shx[3], shy[3], shz[3];            // image shifts (stage 1)
    double K[3];        // stage 1: dt (trmi + trmj); stage 2: (trmi + trmj) |dd|^2
    double ti[3], tj[3];   // trmi, trmj of each constraint
};

// The molecule of this lane, or nGrp for an idle lane.
__device__ inline int m_group(const ShakeArgs &A)
{
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    return g < A.nGrp ? g : A.nGrp;
}

__device__ inline void m_load(const ShakeArgs &A, int g, M3 &M)
{
    const int ao = A.atomOff[g], co = A.conOff[g];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = A.atoms[ao + k];
        M.at[k] = i;
        M.x[k] = A.x[i];
        M.y[k] = A.y[i];
        M.z[k] = A.z[i];
        M.vx[k] = A.vx[i];
        M.vy[k] = A.vy[i];
        M.vz[k] = A.vz[i];
        M.rm[k] = A.rmass[i];
        M.nat[k] = 1.0 / (double)A.nAtConst[i];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int id = A.cons[co + c];
        M.cid[c] = id;
        M.rc2[c] = A.rc2[id];
        M.thr[c] = A.thr[id];
        M.ddx[c] = M.ddy[c] = M.ddz[c] = M.ddsq[c] = 0.;
        M.shx[c] = M.shy[c] = M.shz[c] = 0.;
    }
}

__device__ inline void m_bond(const M3 &M, int c, double &dx, double &dy, double &dz)
{
    dx = M.x[M_B[c]] - M.x[M_A[c]];
    dy = M.y[M_B[c]] - M.y[M_A[c]];
    dz = M.z[M_B[c]] - M.z[M_A[c]];
}

// Constraint vectors at the stage's start (integrate.c:1536-1545).
__device__ inline void m_dd(const ShakeArgs &A, const MdgBox &B, M3 &M)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double dx, dy, dz;
        m_bond(M, c, dx, dy, dz);
        mdg_dist(B, dx, dy, dz);
        M.ddx[c] = dx;
        M.ddy[c] = dy;
        M.ddz[c] = dz;
        A.dd[M.cid[c]] = dx;
        A.dd[A.nConst + M.cid[c]] = dy;
        A.dd[2 * A.nConst + M.cid[c]] = dz;
    }
}

__device__ inline void m_dd_load(const ShakeArgs &A, M3 &M)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        M.ddx[c] = A.dd[M.cid[c]];
        M.ddy[c] = A.dd[A.nConst + M.cid[c]];
        M.ddz[c] = A.dd[2 * A.nConst + M.cid[c]];
    }
}

// Stage 1: trmi = dt rm_a, trmj = dt rm_b, K = dt (trmi + trmj).
__device__ inline void m_setup_r(const ShakeArgs &A, M3 &M)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        M.ti[c] = dt * M.rm[M_A[c]];
        M.tj[c] = dt * M.rm[M_B[c]];
        M.K[c] = dt * (M.ti[c] + M.tj[c]);
        M.ddsq[c] = M.ddx[c] * M.ddx[c] + M.ddy[c] * M.ddy[c] + M.ddz[c] * M.ddz[c];
    }
}

// Stage 2: trmi = 0.5 dt rm_a, trmj = 0.5 dt rm_b, K = (trmi + trmj) |dd|^2.
__device__ inline void m_setup_v(const ShakeArgs &A, M3 &M)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        M.ti[c] = 0.5 * dt * M.rm[M_A[c]];
        M.tj[c] = 0.5 * dt * M.rm[M_B[c]];
        M.K[c] = (M.ti[c] + M.tj[c]) *
                 (M.ddx[c] * M.ddx[c] + M.ddy[c] * M.ddy[c] + M.ddz[c] * M.ddz[c]);
    }
}

__device__ inline void m_kick(const ShakeArgs &A, M3 &M)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = M.at[k];
        M.vx[k] += 0.5 * dt * A.fx[i] * M.rm[k];
        M.vy[k] += 0.5 * dt * A.fy[i] * M.rm[k];
        M.vz[k] += 0.5 * dt * A.fz[i] * M.rm[k];
    }
}

__device__ inline void m_store(const ShakeArgs &A, const M3 &M, bool pos)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = M.at[k];
        if (pos) {
            A.x[i] = M.x[k];
            A.y[i] = M.y[k];
            A.z[i] = M.z[k];
        }
        A.vx[i] = M.vx[k];
        A.vy[i] = M.vy[k];
        A.vz[i] = M.vz[k];
    }
}

__device__ inline void m_scr_store(const ShakeArgs &A, const M3 &M, int slot, bool pos)
{
    const long n = A.P.n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double *s = A.scr + (long)slot * 6 * n + M.at[k];
        if (pos) {
            s[0] = M.x[k];
            s[n] = M.y[k];
            s[2 * n] = M.z[k];
        }
        s[3 * n] = M.vx[k];
        s[4 * n] = M.vy[k];
        s[5 * n] = M.vz[k];
    }
}

__device__ inline void m_scr_load(const ShakeArgs &A, M3 &M, int slot, bool pos)
{
    const long n = A.P.n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double *s = A.scr + (long)slot * 6 * n + M.at[k];
        if (pos) {
            M.x[k] = s[0];
            M.y[k] = s[n];
            M.z[k] = s[2 * n];
        }
        M.vx[k] = s[3 * n];
        M.vy[k] = s[4 * n];
        M.vz[k] = s[5 * n];
    }
}

// The image shift utils.c:38-80 subtracts from each constraint vector, fixed
// for the stage's sweeps (a bond keeps its shift); none for triclinic boxes,
// which take the full image every sweep.
__device__ inline void m_shift_now(const MdgBox &B, M3 &M)
{
#pragma clang fp contract(off)
    if (B.type == 3) return;
    const double Ly = B.type == 1 ? B.a1 : B.b2, Lz = B.type == 1 ? B.a1 : B.c3;
    const double uy = B.type == 1 ? B.u1 : B.v2, uz = B.type == 1 ? B.u1 : B.w3;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double dx, dy, dz;
        m_bond(M, c, dx, dy, dz);
        M.shx[c] = B.a1 * mdg_nint(dx * B.u1);
        M.shy[c] = Ly * mdg_nint(dy * uy);
        M.shz[c] = Lz * mdg_nint(dz * uz);
    }
}

// shake.c:243-258: converged when every constraint has |r^2 - d^2| below
// its threshold; dt*, rt2 keep the vectors for the sweep.
__device__ inline bool m_check(const MdgBox &B, const M3 &M, double (&dtx)[3], double (&dty)[3],
                               double (&dtz)[3], double (&rt2)[3])
{
#pragma clang fp contract(off)
    bool cv = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double dx, dy, dz;
        m_bond(M, c, dx, dy, dz);
        if (B.type == 3) {
            rt2[c] = mdg_dist(B, dx, dy, dz);
        } else {
            dx = dx - M.shx[c];
            dy = dy - M.shy[c];
            dz = dz - M.shz[c];
            rt2[c] = dx * dx + dy * dy + dz * dz;
        }
        dtx[c] = dx;
        dty[c] = dy;
        dtz[c] = dz;
        cv = cv && fabs(rt2[c] - M.rc2[c]) < M.thr[c];
    }
    return cv;
}

// Per-atom sums of the constraint terms t_c (a side ta, b side tb) in
// constraint order from zero, as the reference accumulates xt[].
__device__ inline void m_gather(const double (&ta)[3], const double (&tb)[3], double (&t)[3])
{
#pragma clang fp contract(off)
    t[0] = (0. + ta[0]) + ta[1];
    t[1] = (0. + tb[0]) + ta[2];
    t[2] = (0. + tb[1]) + tb[2];
}

// shake.c:262-332: one position sweep; vir[c] collects constraint c's
// virial terms.
__device__ inline void m_apply_r(const ShakeArgs &A, M3 &M, const double (&dtx)[3],
                                 const double (&dty)[3], const double (&dtz)[3],
                                 const double (&rt2)[3], double (&vir)[3])
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    double ax[3], ay[3], az[3], bx[3], by[3], bz[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double dot = M.ddx[c] * dtx[c] + M.ddy[c] * dty[c] + M.ddz[c] * dtz[c];
        const double lambda = -(M.rc2[c] - rt2[c]) / (M.K[c] * dot);
        vir[c] += lambda * M.ddsq[c];
        const double h = 0.5 * lambda;
        const double li = h * M.ti[c], lj = h * -M.tj[c];
        ax[c] = M.ddx[c] * li;
        ay[c] = M.ddy[c] * li;
        az[c] = M.ddz[c] * li;
        bx[c] = M.ddx[c] * lj;
        by[c] = M.ddy[c] * lj;
        bz[c] = M.ddz[c] * lj;
    }
    double xt[3], yt[3], zt[3];
    m_gather(ax, bx, xt);
    m_gather(ay, by, yt);
    m_gather(az, bz, zt);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        // the correction once per constraint membership (two per atom)
        const double ix = dt * xt[k] * M.nat[k], iy = dt * yt[k] * M.nat[k],
                     iz = dt * zt[k] * M.nat[k];
        M.x[k] = (M.x[k] + ix) + ix;
        M.y[k] = (M.y[k] + iy) + iy;
        M.z[k] = (M.z[k] + iz) + iz;
        const double jx = xt[k] * M.nat[k], jy = yt[k] * M.nat[k], jz = zt[k] * M.nat[k];
        M.vx[k] = (M.vx[k] + jx) + jx;
        M.vy[k] = (M.vy[k] + jy) + jy;
        M.vz[k] = (M.vz[k] + jz) + jz;
    }
}

// shake.c:378-413: one velocity sweep's multipliers; converged when every
// |lambda| is below tolvel.  Per-atom corrections in xt, yt, zt.
__device__ inline bool m_lambda_v(const M3 &M, double tolvel, double (&xt)[3], double (&yt)[3],
                                  double (&zt)[3])
{
#pragma clang fp contract(off)
    double ax[3], ay[3], az[3], bx[3], by[3], bz[3];
    bool cv = true;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = M_A[c], b = M_B[c];
        const double dvx = M.vx[b] - M.vx[a], dvy = M.vy[b] - M.vy[a], dvz = M.vz[b] - M.vz[a];
        const double lambda = (M.ddx[c] * dvx + M.ddy[c] * dvy + M.ddz[c] * dvz) / M.K[c];
        cv = cv && fabs(lambda) < tolvel;
        const double li = lambda * M.ti[c], lj = lambda * -M.tj[c];
        ax[c] = M.ddx[c] * li;
        ay[c] = M.ddy[c] * li;
        az[c] = M.ddz[c] * li;
        bx[c] = M.ddx[c] * lj;
        by[c] = M.ddy[c] * lj;
        bz[c] = M.ddz[c] * lj;
    }
    m_gather(ax, bx, xt);
    m_gather(ay, by, yt);
    m_gather(az, bz, zt);
    return cv;
}

// shake.c:418-436: the corrections, once per constraint membership.
__device__ inline void m_apply_v(M3 &M, const double (&xt)[3], const double (&yt)[3],
                                 const double (&zt)[3])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double jx = xt[k] * M.nat[k], jy = yt[k] * M.nat[k], jz = zt[k] * M.nat[k];
        M.vx[k] = (M.vx[k] + jx) + jx;
        M.vy[k] = (M.vy[k] + jy) + jy;
        M.vz[k] = (M.vz[k] + jz) + jz;
    }
}

// The constraints' running SHAKE virial sums at a saved state.
__device__ inline void m_vir_store(const ShakeArgs &A, int g, int slot, const double (&vir)[3])
{
#pragma unroll
    for (int c = 0; c < 3; c++) A.virScr[((long)slot * A.nGrp + g) * 3 + c] = vir[c];
}

// v + t0 + t1 + t2 in that order.
__device__ inline double m_sum3(const double (&t)[3])
{
#pragma clang fp contract(off)
    return ((0. + t[0]) + t[1]) + t[2];
}

// Converged-sweep bits: the current 64-sweep word in a register, flushed
// when the word is complete or the loop ends.
struct MBits {
    unsigned long long w[SHK_MASKW];
    unsigned long long cur;
};

__device__ inline void mb_flush(MBits &b, int word)
{
#pragma unroll
    for (int q = 0; q < SHK_MASKW; q++)
        if (q == word) b.w[q] = b.cur;
    b.cur = 0;
}

template <bool BATCH>
__device__ inline const MdgBox &m_box(const ShakeArgs &A, int g)
{
    return BATCH ? A.boxes[A.grep[g]] : A.B;
}

template <bool BATCH>
__device__ inline const unsigned long long *m_mask(const ShakeArgs &A, int g)
{
    return BATCH ? A.mask + (long)A.grep[g] * SHK_MASKW : A.mask;
}

template <bool BATCH>
__device__ inline void m_publish(const ShakeArgs &A, int g, const unsigned long long (&bits)[SHK_MASKW],
                                 int kend, bool settled)
{
    if (BATCH) shake_publish_mask_b(A, bits, kend, settled, g < A.nGrp ? A.grep[g] : -1);
    else shake_publish_mask(A, bits, kend, settled);
}

}  // namespace

// Stage 1, part A: half kick, drift, SHAKE-R sweeps of each molecule on its
// own, recording the sweeps at which it was converged and the states of a
// window after its first convergence.  Blocks past S.nbMain do the previous
// step's record and list prune (mdg_side.h).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_m3_vv1_a(ShakeArgs A, SideWork S)
{
#pragma clang fp contract(off)
    if ((int)blockIdx.x >= S.nbMain) {
        side_work<BATCH>(S);
        return;
    }
    const int g = m_group(A);
    // every lane reaches the wave-wide AND; idle lanes contribute all ones
    MBits bits;
#pragma unroll
    for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = ~0ull;
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = m_box<BATCH>(A, g);
        M3 M;
        m_load(A, g, M);
        const double dt = A.P.timeStep;
        m_dd(A, B, M);
        m_setup_r(A, M);
        m_kick(A, M);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            M.x[k] += dt * M.vx[k];
            M.y[k] += dt * M.vy[k];
            M.z[k] += dt * M.vz[k];
        }
        m_shift_now(B, M);
        m_scr_store(A, M, 0, true);
#pragma unroll
        for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = 0;
        bits.cur = 0;
        double vir[3] = {0., 0., 0.};
        int first = -1, it = 0, nr = 0;
        int rr[SHK_R] = {-1, -1, -1, -1};
        // one system: the wavefront sweeps as a whole; a batch (measured
        // slower that way), molecule by molecule
        const bool whole = !BATCH;
        bool prev = false;
        settled = false;
        // molecules sweep on until every molecule of the wavefront has
        // settled, with a ring of the states at wavefront-converged sweeps
        // (see k_w3_vv1_a)
        for (;;) {
            double dtx[3], dty[3], dtz[3], rt2[3];
            const bool cv = m_check(B, M, dtx, dty, dtz, rt2);
            if (cv) bits.cur |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) {
                m_scr_store(A, M, 1 + it - first, true);
                m_vir_store(A, g, 1 + it - first, vir);
            }
            if (whole && __builtin_amdgcn_ballot_w64(!cv) == 0) {
                m_scr_store(A, M, SHK_W + 2 + (nr & (SHK_R - 1)), true);
                m_vir_store(A, g, SHK_W + 2 + (nr & (SHK_R - 1)), vir);
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
            m_apply_r(A, M, dtx, dty, dtz, rt2, vir);
            if ((it & 63) == 63) mb_flush(bits, it >> 6);
            it++;
        }
        mb_flush(bits, it >> 6);
        m_scr_store(A, M, SHK_W + 1, true);
        m_vir_store(A, g, SHK_W + 1, vir);
        A.first[g] = first;
        A.kend[g] = it;
        for (int q = 0; q < SHK_R; q++) A.ring[(long)SHK_R * g + q] = rr[q];
        kend = it;
    }
    m_publish<BATCH>(A, g, bits.w, kend, settled);
}

// Stage 1, part B: every molecule continues (or restarts from a saved
// state) to the first sweep C at which all molecules were converged.  On a
// device step without a list rebuild it also writes the slot positions and
// clears the slot force accumulators (k_w3_vv1_b does the same).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_m3_vv1_b(ShakeArgs A)
{
    // the stage-2 mask (beside this stage's) for this step's stage 2
    if (!BATCH && blockIdx.x == 0 && threadIdx.x < SHK_MASKW) A.mask[SHK_MASKW + threadIdx.x] = ~0ull;
    const int g = m_group(A);
    double vblk = 0.;   // this lane's molecule's SHAKE virial
    double d2 = 0.;     // largest squared move of this lane's atoms
    double rel = 0.;    // largest |relative coordinate| of this lane's atoms
    const bool track = A.track && A.ctrl[CTRL_PRUNE] == 1;
    if (g < A.nGrp) {
        M3 M;
        m_load(A, g, M);
        const int C = shake_global_stop(A, m_mask<BATCH>(A, g));
        if (C < 0) {
            atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 1);
        } else {
            const MdgBox &B = m_box<BATCH>(A, g);
            const int first = A.first[g], kend = A.kend[g];
            int slot, at;
            shake_pick_slot(C, first, kend, A.ring + (long)SHK_R * g, slot, at);
            m_dd_load(A, M);
            m_setup_r(A, M);
            m_scr_load(A, M, slot, true);
            m_shift_now(B, M);
            // each constraint continues its own running sum, as in the quad
            // kernels
            double vir[3] = {0., 0., 0.};
            if (slot != 0)
                for (int c = 0; c < 3; c++) vir[c] = A.virScr[((long)slot * A.nGrp + g) * 3 + c];
            for (int it = at; it < C; it++) {
                double dtx[3], dty[3], dtz[3], rt2[3];
                m_check(B, M, dtx, dty, dtz, rt2);
                m_apply_r(A, M, dtx, dty, dtz, rt2, vir);
            }
            {
                // the state at C passes the check, unless a molecule
                // re-diverged after it was taken as settled
                double dtx[3], dty[3], dtz[3], rt2[3];
                if (!m_check(B, M, dtx, dty, dtz, rt2)) atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
            }
            const double vm = m_sum3(vir);
            if (BATCH) A.vir[g] = vm;   // per molecule, summed per replica
            else vblk = vm;
            m_store(A, M, true);
        }
        if (!BATCH && A.slotWrite) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int i = M.at[k];
                const int s = A.atomSlot[i];
                const double nx = M.x[k], ny = M.y[k], nz = M.z[k];   // as stored above
                if (track) {
                    const double L0 = A.B.a1, L1 = A.B.type == 1 ? A.B.a1 : A.B.b2,
                                 L2 = A.B.type == 1 ? A.B.a1 : A.B.c3;
                    double dx = nx - A.sx[s], dy = ny - A.sy[s], dz = nz - A.sz[s];
                    dx -= L0 * rint(dx / L0);
                    dy -= L1 * rint(dy / L1);
                    dz -= L2 * rint(dz / L2);
                    double e = dx * dx + dy * dy + dz * dz;
                    if (!(e <= 1e300)) e = INFINITY;   // NaN positions: no pruning
                    d2 = fmax(d2, e);
                }
                if (A.spr)
                    rel = fmax(rel, mdg_slot_rel(A.B, A.cO, A.ncluO, s / MDG_CL, nx, ny, nz,
                                                 A.spr, s));
                A.sx[s] = nx;
                A.sy[s] = ny;
                A.sz[s] = nz;
                A.fix[s] = 0;
                A.fix[s + A.fstr] = 0;
                A.fix[s + 2 * A.fstr] = 0;
            }
        }
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
// same bookkeeping as stage 1.  With fuse set (1 mixed, 2 FP64; one box) the
// kernel first assembles each atom's force (fin_atom, mdg_side.h), keeps the
// correction partials per block and resets the stage-1 SHAKE mask for the
// next step, as k_w3_vv2_a does.
template <bool BATCH>
__global__ void __launch_bounds__(256) k_m3_vv2_a(ShakeArgs A, FinArgs F, int fuse)
{
    __shared__ double mean[3], red[4][3];
    double facc[2] = {0., 0.};
    if (!BATCH && fuse) {
        fin_mean(F, mean, red);
        if (blockIdx.x == 0 && threadIdx.x < SHK_MASKW) F.mask[threadIdx.x] = ~0ull;
    }
    const int g = m_group(A);
    MBits bits;
#pragma unroll
    for (int w = 0; w < SHK_MASKW; w++) bits.w[w] = ~0ull;
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) {
        const MdgBox &B = m_box<BATCH>(A, g);
        M3 M;
        m_load(A, g, M);
        if (!BATCH && fuse) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double gx, gy, gz;
                const bool fr = fuse == 1 ? fin_atom<true>(F, M.at[k], mean, facc, gx, gy, gz)
                                          : fin_atom<false>(F, M.at[k], mean, facc, gx, gy, gz);
                if (fr) M.vx[k] = M.vy[k] = M.vz[k] = 0.;
            }
        }
        m_dd(A, B, M);
        m_setup_v(A, M);
        m_kick(A, M);
        m_scr_store(A, M, 0, false);
        const double tolvel = A.P.tolShake * A.P.rTimeStep;
#pragma unroll
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
            // reference's global loop goes on while any group is unconverged
            double xt[3], yt[3], zt[3];
            const bool cv = m_lambda_v(M, tolvel, xt, yt, zt);
            if (cv) bits.cur |= 1ull << (it & 63);
            if (cv && first < 0) first = it;
            if (first >= 0 && it - first < SHK_W) m_scr_store(A, M, 1 + it - first, false);
            if (whole && __builtin_amdgcn_ballot_w64(!cv) == 0) {
                m_scr_store(A, M, SHK_W + 2 + (nr & (SHK_R - 1)), false);
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
            m_apply_v(M, xt, yt, zt);
            if ((it & 63) == 63) mb_flush(bits, it >> 6);
            it++;
        }
        mb_flush(bits, it >> 6);
        m_scr_store(A, M, SHK_W + 1, false);
        A.first[g] = first;
        A.kend[g] = it;
        for (int q = 0; q < SHK_R; q++) A.ring[(long)SHK_R * g + q] = rr[q];
        kend = it;
    }
    m_publish<BATCH>(A, g, bits.w, kend, settled);
    if (!BATCH && fuse) mdg_block_partials<2>(facc, F.part);
}

// Stage 2, part B: replay to the global sweep C, kinetic energy and stress
// partials (utils.c:502-515), wrap (utils.c:162-218).
template <bool BATCH>
__global__ void __launch_bounds__(256) k_m3_vv2_b(ShakeArgs A)
{
#pragma clang fp contract(off)
    const int g = m_group(A);
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (g < A.nGrp) {
        const MdgBox &B = m_box<BATCH>(A, g);
        M3 M;
        m_load(A, g, M);
        const int C = shake_global_stop(A, m_mask<BATCH>(A, g));
        if (C < 0) {
            atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 2);
        } else {
            const int first = A.first[g], kend = A.kend[g];
            int slot, at;
            shake_pick_slot(C, first, kend, A.ring + (long)SHK_R * g, slot, at);
            m_dd_load(A, M);
            m_setup_v(A, M);
            m_scr_load(A, M, slot, false);
            for (int it = at; it < C; it++) {
                double xt[3], yt[3], zt[3];
                m_lambda_v(M, 0., xt, yt, zt);
                m_apply_v(M, xt, yt, zt);
            }
            {
                double xt[3], yt[3], zt[3];
                if (!m_lambda_v(M, A.P.tolShake * A.P.rTimeStep, xt, yt, zt))
                    atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
            }
        }
        double ke[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m = A.mass[M.at[k]];
            const double vx = M.vx[k], vy = M.vy[k], vz = M.vz[k];
            ke[k] = m * (vx * vx + vy * vy + vz * vz);
            acc[0] += ke[k];
            acc[1] += m * vx * vx;
            acc[2] += m * vx * vy;
            acc[3] += m * vx * vz;
            acc[4] += m * vy * vy;
            acc[5] += m * vy * vz;
            acc[6] += m * vz * vz;
            shake_image(B, M.x[k], M.y[k], M.z[k]);
        }
        m_store(A, M, true);
        if (BATCH) A.grpKin[g] = m_sum3(ke);   // the molecule's sum, in atom order
    }
    mdg_block_partials<7>(acc, A.part);
}

template __global__ void k_m3_vv1_a<false>(ShakeArgs, SideWork);
template __global__ void k_m3_vv1_a<true>(ShakeArgs, SideWork);
template __global__ void k_m3_vv1_b<false>(ShakeArgs);
template __global__ void k_m3_vv1_b<true>(ShakeArgs);
template __global__ void k_m3_vv2_a<false>(ShakeArgs, FinArgs, int);
template __global__ void k_m3_vv2_a<true>(ShakeArgs, FinArgs, int);
template __global__ void k_m3_vv2_b<false>(ShakeArgs);
template __global__ void k_m3_vv2_b<true>(ShakeArgs);
