// mdg_shake_w3.hip -- VV + SHAKE kernels specialised for rigid
// three-site molecules (three atoms a0 < a1 < a2 with constraints
// (a0,a1), (a0,a2), (a1,a2) in that order, e.g. TIP3P water under
// consth).  Same arithmetic, operation order and global stopping rule as
// the generic kernels of mdg_integrate.hip; every index is a compile-time
// constant, so the molecule's state lives in registers.
#include "mdg_shake.h"

namespace {

constexpr int IA[3] = {0, 0, 1};
constexpr int IB[3] = {1, 2, 2};

struct W3 {
    int at[3];
    int cid[3];
    double x[3], y[3], z[3], vx[3], vy[3], vz[3];
    double rm[3], nat[3];
    double ddx[3], ddy[3], ddz[3], rc2[3], src2[3];
};

__device__ inline void w3_load(const ShakeArgs &A, int g, W3 &G)
{
    const int a0 = A.atomOff[g], c0 = A.conOff[g];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = A.atoms[a0 + k];
        G.at[k] = i;
        G.x[k] = A.x[i];
        G.y[k] = A.y[i];
        G.z[k] = A.z[i];
        G.vx[k] = A.vx[i];
        G.vy[k] = A.vy[i];
        G.vz[k] = A.vz[i];
        G.rm[k] = A.rmass[i];
        G.nat[k] = 1.0 / (double)A.nAtConst[i];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int ci = A.cons[c0 + c];
        G.cid[c] = ci;
        G.rc2[c] = A.rc2[ci];
        G.src2[c] = sqrt(G.rc2[c]);
    }
}

__device__ inline void w3_store(const ShakeArgs &A, const W3 &G, bool pos)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = G.at[k];
        if (pos) {
            A.x[i] = G.x[k];
            A.y[i] = G.y[k];
            A.z[i] = G.z[k];
        }
        A.vx[i] = G.vx[k];
        A.vy[i] = G.vy[k];
        A.vz[i] = G.vz[k];
    }
}

__device__ inline void w3_scr_store(const ShakeArgs &A, const W3 &G, int slot, bool pos)
{
    const long n = A.P.n;
    double *s = A.scr + (long)slot * 6 * n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = G.at[k];
        if (pos) {
            s[i] = G.x[k];
            s[n + i] = G.y[k];
            s[2 * n + i] = G.z[k];
        }
        s[3 * n + i] = G.vx[k];
        s[4 * n + i] = G.vy[k];
        s[5 * n + i] = G.vz[k];
    }
}

__device__ inline void w3_scr_load(const ShakeArgs &A, W3 &G, int slot, bool pos)
{
    const long n = A.P.n;
    const double *s = A.scr + (long)slot * 6 * n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = G.at[k];
        if (pos) {
            G.x[k] = s[i];
            G.y[k] = s[n + i];
            G.z[k] = s[2 * n + i];
        }
        G.vx[k] = s[3 * n + i];
        G.vy[k] = s[4 * n + i];
        G.vz[k] = s[5 * n + i];
    }
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

__device__ inline void w3_dd(const ShakeArgs &A, const MdgBox &B, W3 &G)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double dx = G.x[IB[c]] - G.x[IA[c]];
        double dy = G.y[IB[c]] - G.y[IA[c]];
        double dz = G.z[IB[c]] - G.z[IA[c]];
        w3_dist(B, dx, dy, dz);
        G.ddx[c] = dx;
        G.ddy[c] = dy;
        G.ddz[c] = dz;
        A.dd[G.cid[c]] = dx;
        A.dd[A.nConst + G.cid[c]] = dy;
        A.dd[2 * A.nConst + G.cid[c]] = dz;
    }
}

__device__ inline void w3_dd_load(const ShakeArgs &A, W3 &G)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        G.ddx[c] = A.dd[G.cid[c]];
        G.ddy[c] = A.dd[A.nConst + G.cid[c]];
        G.ddz[c] = A.dd[2 * A.nConst + G.cid[c]];
    }
}

// shake.c:243-258
__device__ inline double w3_check(const ShakeArgs &A, const MdgBox &B, const W3 &G, double *dtx,
                                  double *dty, double *dtz, double *rt2)
{
#pragma clang fp contract(off)
    double maxd = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        dtx[c] = G.x[IB[c]] - G.x[IA[c]];
        dty[c] = G.y[IB[c]] - G.y[IA[c]];
        dtz[c] = G.z[IB[c]] - G.z[IA[c]];
        rt2[c] = w3_dist(B, dtx[c], dty[c], dtz[c]);
        const double d = fabs(rt2[c] - G.rc2[c]) / G.src2[c];
        maxd = maxd >= d ? maxd : d;
    }
    return 0.5 * maxd;
}

// shake.c:262-332
__device__ inline void w3_apply_r(const ShakeArgs &A, W3 &G, const double *dtx, const double *dty,
                                  const double *dtz, const double *rt2, double &vir)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    double xt[3] = {0., 0., 0.}, yt[3] = {0., 0., 0.}, zt[3] = {0., 0., 0.};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = IA[c], b = IB[c];
        const double trmi = dt * G.rm[a], trmj = dt * G.rm[b];
        const double lambda = -(G.rc2[c] - rt2[c]) /
            (dt * (trmi + trmj) * ((G.ddx[c] * dtx[c]) + (G.ddy[c] * dty[c]) + (G.ddz[c] * dtz[c])));
        vir += lambda * (G.ddx[c] * G.ddx[c] + G.ddy[c] * G.ddy[c] + G.ddz[c] * G.ddz[c]);
        const double li = 0.5 * lambda * trmi;
        xt[a] += G.ddx[c] * li;
        yt[a] += G.ddy[c] * li;
        zt[a] += G.ddz[c] * li;
        const double lj = -0.5 * lambda * trmj;
        xt[b] += G.ddx[c] * lj;
        yt[b] += G.ddy[c] * lj;
        zt[b] += G.ddz[c] * lj;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = IA[c], b = IB[c];
        G.x[a] += dt * xt[a] * G.nat[a];
        G.y[a] += dt * yt[a] * G.nat[a];
        G.z[a] += dt * zt[a] * G.nat[a];
        G.x[b] += dt * xt[b] * G.nat[b];
        G.y[b] += dt * yt[b] * G.nat[b];
        G.z[b] += dt * zt[b] * G.nat[b];
        G.vx[a] += xt[a] * G.nat[a];
        G.vy[a] += yt[a] * G.nat[a];
        G.vz[a] += zt[a] * G.nat[a];
        G.vx[b] += xt[b] * G.nat[b];
        G.vy[b] += yt[b] * G.nat[b];
        G.vz[b] += zt[b] * G.nat[b];
    }
}

// shake.c:378-413: velocity lambdas of one sweep; returns max |lambda|
// and the per-atom corrections.
__device__ inline double w3_lambda_v(const ShakeArgs &A, const W3 &G, const double *den,
                                     double *xt, double *yt, double *zt)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    double maxd = 0.;
#pragma unroll
    for (int k = 0; k < 3; k++) xt[k] = yt[k] = zt[k] = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = IA[c], b = IB[c];
        const double trmi = 0.5 * dt * G.rm[a], trmj = 0.5 * dt * G.rm[b];
        const double lambda = (G.ddx[c] * (G.vx[b] - G.vx[a]) + G.ddy[c] * (G.vy[b] - G.vy[a]) +
                               G.ddz[c] * (G.vz[b] - G.vz[a])) / den[c];
        const double al = fabs(lambda);
        maxd = maxd >= al ? maxd : al;
        const double li = lambda * trmi;
        xt[a] += G.ddx[c] * li;
        yt[a] += G.ddy[c] * li;
        zt[a] += G.ddz[c] * li;
        const double lj = -lambda * trmj;
        xt[b] += G.ddx[c] * lj;
        yt[b] += G.ddy[c] * lj;
        zt[b] += G.ddz[c] * lj;
    }
    return maxd;
}

// shake.c:418-436: apply one sweep's corrections, once per constraint
// membership with weight 1/nAtConst.
__device__ inline void w3_apply_v(W3 &G, const double *xt, const double *yt, const double *zt)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = IA[c], b = IB[c];
        G.vx[a] += xt[a] * G.nat[a];
        G.vy[a] += yt[a] * G.nat[a];
        G.vz[a] += zt[a] * G.nat[a];
        G.vx[b] += xt[b] * G.nat[b];
        G.vy[b] += yt[b] * G.nat[b];
        G.vz[b] += zt[b] * G.nat[b];
    }
}

// (trmi + trmj) * |dd|^2, the velocity lambda's denominator (shake.c:391-393)
__device__ inline void w3_den(const ShakeArgs &A, const W3 &G, double *den)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double trmi = 0.5 * dt * G.rm[IA[c]], trmj = 0.5 * dt * G.rm[IB[c]];
        den[c] = (trmi + trmj) * (G.ddx[c] * G.ddx[c] + G.ddy[c] * G.ddy[c] + G.ddz[c] * G.ddz[c]);
    }
}

__device__ inline void w3_kick(const ShakeArgs &A, W3 &G)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = G.at[k];
        G.vx[k] += 0.5 * dt * A.fx[i] * G.rm[k];
        G.vy[k] += 0.5 * dt * A.fy[i] * G.rm[k];
        G.vz[k] += 0.5 * dt * A.fz[i] * G.rm[k];
    }
}

}  // namespace

__device__ inline void w3_vv1_a_body(const ShakeArgs &A, const MdgBox &B, int g,
                                     unsigned long long (&bits)[SHK_MASKW], int &kend, bool &settled)
{
#pragma clang fp contract(off)
    W3 G;
    w3_load(A, g, G);
    const double dt = A.P.timeStep;
    w3_dd(A, B, G);
    w3_kick(A, G);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        G.x[k] += dt * G.vx[k];
        G.y[k] += dt * G.vy[k];
        G.z[k] += dt * G.vz[k];
    }
    w3_scr_store(A, G, 0, true);
    for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
    double dtx[3], dty[3], dtz[3], rt2[3];
    double vir = 0.;
    int first = -1, k = 0;
    bool prev = false;
    settled = false;
    for (;;) {
        const bool cv = w3_check(A, B, G, dtx, dty, dtz, rt2) < A.P.tolShake;
        if (cv) bits[k >> 6] |= 1ull << (k & 63);
        if (cv && first < 0) first = k;
        if (first >= 0 && k - first < SHK_W) {
            w3_scr_store(A, G, 1 + k - first, true);
            A.virScr[(long)(1 + k - first) * A.nGrp + g] = vir;
        }
        if (cv && prev) {
            settled = true;
            break;
        }
        prev = cv;
        if (k + 1 >= A.P.maxCycle) break;
        w3_apply_r(A, G, dtx, dty, dtz, rt2, vir);
        k++;
    }
    w3_scr_store(A, G, SHK_W + 1, true);
    A.virScr[(long)(SHK_W + 1) * A.nGrp + g] = vir;
    A.first[g] = first;
    A.kend[g] = k;
    kend = k;
}

// The box of group g: the run's, or its replica's in a batch.
template <bool BATCH>
__device__ inline const MdgBox &w3_box(const ShakeArgs &A, int g)
{
    return BATCH ? A.boxes[A.grep[g]] : A.B;
}

template <bool BATCH>
__device__ inline void w3_publish(const ShakeArgs &A, int g, const unsigned long long (&bits)[SHK_MASKW],
                                  int kend, bool settled)
{
    if (BATCH) shake_publish_mask_b(A, bits, kend, settled, g < A.nGrp ? A.grep[g] : -1);
    else shake_publish_mask(A, bits, kend, settled);
}

template <bool BATCH>
__device__ inline const unsigned long long *w3_mask(const ShakeArgs &A, int g)
{
    return BATCH ? A.mask + (long)A.grep[g] * SHK_MASKW : A.mask;
}

template <bool BATCH>
__global__ void __launch_bounds__(64) k_w3_vv1_a(ShakeArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the wave-wide AND; idle lanes contribute all ones
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) w3_vv1_a_body(A, w3_box<BATCH>(A, g), g, bits, kend, settled);
    w3_publish<BATCH>(A, g, bits, kend, settled);
}

template <bool BATCH>
__global__ void __launch_bounds__(64) k_w3_vv1_b(ShakeArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nGrp) return;
    const MdgBox &B = w3_box<BATCH>(A, g);
    const int C = shake_global_stop(A, w3_mask<BATCH>(A, g));
    if (C < 0) {
        atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 1);
        return;
    }
    const int first = A.first[g], kend = A.kend[g];
    int slot, at;
    shake_pick_slot(C, first, kend, slot, at);
    W3 G;
    w3_load(A, g, G);
    w3_dd_load(A, G);
    w3_scr_load(A, G, slot, true);
    double vir = slot == 0 ? 0. : A.virScr[(long)slot * A.nGrp + g];
    double dtx[3], dty[3], dtz[3], rt2[3];
    for (int k = at; k < C; k++) {
        w3_check(A, B, G, dtx, dty, dtz, rt2);
        w3_apply_r(A, G, dtx, dty, dtz, rt2, vir);
    }
    A.vir[g] = vir;
    w3_store(A, G, true);
}

__device__ inline void w3_vv2_a_body(const ShakeArgs &A, const MdgBox &B, int g,
                                     unsigned long long (&bits)[SHK_MASKW], int &kend, bool &settled)
{
    W3 G;
    w3_load(A, g, G);
    w3_dd(A, B, G);
    w3_kick(A, G);
    w3_scr_store(A, G, 0, false);
    double den[3];
    w3_den(A, G, den);
    const double tolvel = A.P.tolShake * A.P.rTimeStep;
    for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
    int first = -1, k = 0;
    bool prev = false;
    settled = false;
    double xt[3], yt[3], zt[3];
    for (;;) {
        // the check is the sweep's own max |lambda| (shake.c:421); every
        // sweep that is not the last one checked is applied, because the
        // reference's global loop goes on while any group is unconverged
        const bool cv = w3_lambda_v(A, G, den, xt, yt, zt) < tolvel;
        if (cv) bits[k >> 6] |= 1ull << (k & 63);
        if (cv && first < 0) first = k;
        if (first >= 0 && k - first < SHK_W) w3_scr_store(A, G, 1 + k - first, false);
        if (cv && prev) {
            settled = true;
            break;
        }
        prev = cv;
        if (k + 1 >= A.P.maxCycle) break;
        w3_apply_v(G, xt, yt, zt);
        k++;
    }
    w3_scr_store(A, G, SHK_W + 1, false);
    A.first[g] = first;
    A.kend[g] = k;
    kend = k;
}

template <bool BATCH>
__global__ void __launch_bounds__(64) k_w3_vv2_a(ShakeArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the wave-wide AND; idle lanes contribute all ones
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) w3_vv2_a_body(A, w3_box<BATCH>(A, g), g, bits, kend, settled);
    w3_publish<BATCH>(A, g, bits, kend, settled);
}

template <bool BATCH>
__global__ void __launch_bounds__(64) k_w3_vv2_b(ShakeArgs A)
{
#pragma clang fp contract(off)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (g < A.nGrp) {
        const MdgBox &B = w3_box<BATCH>(A, g);
        W3 G;
        w3_load(A, g, G);
        const int C = shake_global_stop(A, w3_mask<BATCH>(A, g));
        if (C < 0) {
            atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 2);
        } else {
            const int first = A.first[g], kend = A.kend[g];
            int slot, at;
            shake_pick_slot(C, first, kend, slot, at);
            w3_dd_load(A, G);
            w3_scr_load(A, G, slot, false);
            double den[3];
            w3_den(A, G, den);
            double xt[3], yt[3], zt[3];
            for (int k = at; k < C; k++) {
                w3_lambda_v(A, G, den, xt, yt, zt);
                w3_apply_v(G, xt, yt, zt);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double m = A.mass[G.at[k]];
            acc[0] += m * (G.vx[k] * G.vx[k] + G.vy[k] * G.vy[k] + G.vz[k] * G.vz[k]);
            acc[1] += m * G.vx[k] * G.vx[k];
            acc[2] += m * G.vx[k] * G.vy[k];
            acc[3] += m * G.vx[k] * G.vz[k];
            acc[4] += m * G.vy[k] * G.vy[k];
            acc[5] += m * G.vy[k] * G.vz[k];
            acc[6] += m * G.vz[k] * G.vz[k];
            shake_image(B, G.x[k], G.y[k], G.z[k]);
        }
        w3_store(A, G, true);
        if (BATCH) A.grpKin[g] = acc[0];
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
