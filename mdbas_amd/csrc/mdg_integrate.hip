// mdg_integrate.hip -- velocity Verlet NVE with SHAKE on the GPU
// (integrate.c:1518-1634, shake.c:209-457).
//
// The reference SHAKE is a Jacobi sweep over all constraints that stops
// at the first sweep index C at which the largest residual of the whole
// system is below tol.  Constraint groups (connected components of the
// constraint graph; a water molecule is one group) never share atoms, so a
// group's state after k sweeps does not depend on other groups; only C is
// global.  A group's residual can dip below tol and rise above it again
// (the averaged Jacobi iteration oscillates with period two), so C is not
// simply the largest per-group first-convergence index.  Kernel A runs
// each group (one thread) until it has been converged at two consecutive
// checks, records the converged / not-converged bit of every sweep index
// it checked, ANDs those bits into a device-wide mask and keeps its states
// for the first SHK_W sweeps from its first convergence on.  C is the first
// set bit of the mask; kernel B gives every group its state after exactly
// C sweeps (from the kept window, or by sweeping on), as the reference's
// global loop would, with no host round trip.  A group settled at two
// consecutive checks is taken as converged from then on; an oscillation
// wider than that would be a divergence from the reference in its SHAKE
// stopping sweep only.  Arithmetic follows the reference's operation order
// with contraction off.
#include "mdg_shake.h"



__device__ inline int local_of(const int *atoms, int na, int g)
{
    for (int k = 0; k < na; k++)
        if (atoms[k] == g) return k;
    return 0;
}

struct Grp {
    int na, nc;
    int at[MDG_GRP_ATOMS];
    int ia[MDG_GRP_CONS], ib[MDG_GRP_CONS], cid[MDG_GRP_CONS];
    double x[MDG_GRP_ATOMS], y[MDG_GRP_ATOMS], z[MDG_GRP_ATOMS];
    double vx[MDG_GRP_ATOMS], vy[MDG_GRP_ATOMS], vz[MDG_GRP_ATOMS];
    double rm[MDG_GRP_ATOMS], nat[MDG_GRP_ATOMS];
    double ddx[MDG_GRP_CONS], ddy[MDG_GRP_CONS], ddz[MDG_GRP_CONS], rc2[MDG_GRP_CONS];
};

__device__ inline void grp_load(const ShakeArgs &A, int g, Grp &G)
{
    const int a0 = A.atomOff[g], c0 = A.conOff[g];
    G.na = A.atomOff[g + 1] - a0;
    G.nc = A.conOff[g + 1] - c0;
    for (int k = 0; k < G.na; k++) {
        int i = A.atoms[a0 + k];
        G.at[k] = i;
        G.x[k] = A.x[i];
        G.y[k] = A.y[i];
        G.z[k] = A.z[i];
        G.vx[k] = A.vx[i];
        G.vy[k] = A.vy[i];
        G.vz[k] = A.vz[i];
        G.rm[k] = A.rmass[i];
        int nc = A.nAtConst[i];
        G.nat[k] = nc > 0 ? 1.0 / (double)nc : 0.0;
    }
    for (int c = 0; c < G.nc; c++) {
        int ci = A.cons[c0 + c];
        G.cid[c] = ci;
        G.ia[c] = local_of(G.at, G.na, A.ca[ci]);
        G.ib[c] = local_of(G.at, G.na, A.cb[ci]);
        G.rc2[c] = A.rc2[ci];
    }
}

__device__ inline void grp_store(const ShakeArgs &A, const Grp &G, bool pos)
{
    for (int k = 0; k < G.na; k++) {
        int i = G.at[k];
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

// shake.c:243-258: residual check over the group; fills dt / rt2.
__device__ inline double shake_r_check(const ShakeArgs &A, const Grp &G, double *dtx, double *dty,
                                       double *dtz, double *rt2)
{
#pragma clang fp contract(off)
    double maxd = 0.;
    for (int c = 0; c < G.nc; c++) {
        dtx[c] = G.x[G.ib[c]] - G.x[G.ia[c]];
        dty[c] = G.y[G.ib[c]] - G.y[G.ia[c]];
        dtz[c] = G.z[G.ib[c]] - G.z[G.ia[c]];
        rt2[c] = mdg_dist(A.B, dtx[c], dty[c], dtz[c]);
        double d = fabs(rt2[c] - G.rc2[c]) / sqrt(G.rc2[c]);
        maxd = maxd >= d ? maxd : d;
    }
    return 0.5 * maxd;
}

// shake.c:262-332: one correction sweep over the group.
__device__ inline void shake_r_apply(const ShakeArgs &A, Grp &G, const double *dtx,
                                     const double *dty, const double *dtz, const double *rt2,
                                     double &vir)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    double xt[MDG_GRP_ATOMS], yt[MDG_GRP_ATOMS], zt[MDG_GRP_ATOMS];
    for (int k = 0; k < G.na; k++) xt[k] = yt[k] = zt[k] = 0.;
    for (int c = 0; c < G.nc; c++) {
        const int a = G.ia[c], b = G.ib[c];
        double trmi = dt * G.rm[a], trmj = dt * G.rm[b];
        double lambda = -(G.rc2[c] - rt2[c]) /
            (dt * (trmi + trmj) * ((G.ddx[c] * dtx[c]) + (G.ddy[c] * dty[c]) + (G.ddz[c] * dtz[c])));
        vir += lambda * (G.ddx[c] * G.ddx[c] + G.ddy[c] * G.ddy[c] + G.ddz[c] * G.ddz[c]);
        double li = 0.5 * lambda * trmi;
        xt[a] += G.ddx[c] * li;
        yt[a] += G.ddy[c] * li;
        zt[a] += G.ddz[c] * li;
        double lj = -0.5 * lambda * trmj;
        xt[b] += G.ddx[c] * lj;
        yt[b] += G.ddy[c] * lj;
        zt[b] += G.ddz[c] * lj;
    }
    for (int c = 0; c < G.nc; c++) {
        const int a = G.ia[c], b = G.ib[c];
        double nia = G.nat[a], nib = G.nat[b];
        G.x[a] += dt * xt[a] * nia;
        G.y[a] += dt * yt[a] * nia;
        G.z[a] += dt * zt[a] * nia;
        G.x[b] += dt * xt[b] * nib;
        G.y[b] += dt * yt[b] * nib;
        G.z[b] += dt * zt[b] * nib;
        G.vx[a] += xt[a] * nia;
        G.vy[a] += yt[a] * nia;
        G.vz[a] += zt[a] * nia;
        G.vx[b] += xt[b] * nib;
        G.vy[b] += yt[b] * nib;
        G.vz[b] += zt[b] * nib;
    }
}

// shake.c:378-430: velocity residual (max |lambda|) and the corrections.
__device__ inline double shake_v_sweep(const ShakeArgs &A, Grp &G, bool apply_if_needed,
                                       double tolvel, bool force_apply)
{
#pragma clang fp contract(off)
    const double dt = A.P.timeStep;
    double xt[MDG_GRP_ATOMS], yt[MDG_GRP_ATOMS], zt[MDG_GRP_ATOMS];
    for (int k = 0; k < G.na; k++) xt[k] = yt[k] = zt[k] = 0.;
    double maxd = 0.;
    for (int c = 0; c < G.nc; c++) {
        const int a = G.ia[c], b = G.ib[c];
        double trmi = 0.5 * dt * G.rm[a], trmj = 0.5 * dt * G.rm[b];
        double lambda = (G.ddx[c] * (G.vx[b] - G.vx[a]) + G.ddy[c] * (G.vy[b] - G.vy[a]) +
                         G.ddz[c] * (G.vz[b] - G.vz[a])) /
                        ((trmi + trmj) *
                         (G.ddx[c] * G.ddx[c] + G.ddy[c] * G.ddy[c] + G.ddz[c] * G.ddz[c]));
        double al = fabs(lambda);
        maxd = maxd >= al ? maxd : al;
        double li = lambda * trmi;
        xt[a] += G.ddx[c] * li;
        yt[a] += G.ddy[c] * li;
        zt[a] += G.ddz[c] * li;
        double lj = -lambda * trmj;
        xt[b] += G.ddx[c] * lj;
        yt[b] += G.ddy[c] * lj;
        zt[b] += G.ddz[c] * lj;
    }
    bool apply = force_apply || (apply_if_needed && !(maxd < tolvel));
    if (apply) {
        for (int c = 0; c < G.nc; c++) {
            const int a = G.ia[c], b = G.ib[c];
            double nia = G.nat[a], nib = G.nat[b];
            G.vx[a] += xt[a] * nia;
            G.vy[a] += yt[a] * nia;
            G.vz[a] += zt[a] * nia;
            G.vx[b] += xt[b] * nib;
            G.vy[b] += yt[b] * nib;
            G.vz[b] += zt[b] * nib;
        }
    }
    return maxd;
}

__device__ inline void scr_store(const ShakeArgs &A, const Grp &G, int slot, bool pos)
{
    const long n = A.P.n;
    double *s = A.scr + (long)slot * 6 * n;
    for (int k = 0; k < G.na; k++) {
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

__device__ inline void scr_load(const ShakeArgs &A, Grp &G, int slot, bool pos)
{
    const long n = A.P.n;
    const double *s = A.scr + (long)slot * 6 * n;
    for (int k = 0; k < G.na; k++) {
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

__device__ inline void dd_load(const ShakeArgs &A, Grp &G)
{
    for (int c = 0; c < G.nc; c++) {
        G.ddx[c] = A.dd[G.cid[c]];
        G.ddy[c] = A.dd[A.nConst + G.cid[c]];
        G.ddz[c] = A.dd[2 * A.nConst + G.cid[c]];
    }
}

__device__ inline void dd_compute(const ShakeArgs &A, Grp &G)
{
    for (int c = 0; c < G.nc; c++) {
        double dx = G.x[G.ib[c]] - G.x[G.ia[c]];
        double dy = G.y[G.ib[c]] - G.y[G.ia[c]];
        double dz = G.z[G.ib[c]] - G.z[G.ia[c]];
        mdg_dist(A.B, dx, dy, dz);
        G.ddx[c] = dx;
        G.ddy[c] = dy;
        G.ddz[c] = dz;
        A.dd[G.cid[c]] = dx;
        A.dd[A.nConst + G.cid[c]] = dy;
        A.dd[2 * A.nConst + G.cid[c]] = dz;
    }
}

// Stage 1, part A: constraint vectors, half kick, drift, then SHAKE-R
// sweeps until the group has been converged at two consecutive checks
// (integrate.c:1526-1590, shake.c:243-338), recording which sweep indices
// were converged and keeping the states of the first SHK_W sweeps from
// the first convergence on.
__device__ inline void vv1_a_body(const ShakeArgs &A, int g, unsigned long long (&bits)[SHK_MASKW],
                                  int &kend, bool &settled)
{
#pragma clang fp contract(off)
    Grp G;
    grp_load(A, g, G);
    const double dt = A.P.timeStep;
    if (A.ddStored)
        dd_load(A, G);
    else
        dd_compute(A, G);
    if (!A.noKick)
        for (int k = 0; k < G.na; k++) {
            int i = G.at[k];
            G.vx[k] += 0.5 * dt * A.fx[i] * G.rm[k];
            G.vy[k] += 0.5 * dt * A.fy[i] * G.rm[k];
            G.vz[k] += 0.5 * dt * A.fz[i] * G.rm[k];
        }
    if (A.npt) {
        // integrate.c:1876-1878 (com zero), 2398-2400
        for (int k = 0; k < G.na; k++) {
            G.x[k] = A.dscale * (G.x[k] - A.com0) + dt * G.vx[k] + A.com0;
            G.y[k] = A.dscale * (G.y[k] - A.com1) + dt * G.vy[k] + A.com1;
            G.z[k] = A.dscale * (G.z[k] - A.com2) + dt * G.vz[k] + A.com2;
        }
    } else {
        for (int k = 0; k < G.na; k++) {
            G.x[k] += dt * G.vx[k];
            G.y[k] += dt * G.vy[k];
            G.z[k] += dt * G.vz[k];
        }
    }
    if (G.nc == 0) {
        grp_store(A, G, true);
        for (int w = 0; w < SHK_MASKW; w++) bits[w] = ~0ull;
        kend = -1;
        settled = true;
        return;
    }
    scr_store(A, G, 0, true);
    for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
    double dtx[MDG_GRP_CONS], dty[MDG_GRP_CONS], dtz[MDG_GRP_CONS], rt2[MDG_GRP_CONS];
    double vir = 0.;
    int first = -1, k = 0;
    bool prev = false;
    settled = false;
    for (;;) {
        const bool cv = shake_r_check(A, G, dtx, dty, dtz, rt2) < A.P.tolShake;
        if (cv) bits[k >> 6] |= 1ull << (k & 63);
        if (cv && first < 0) first = k;
        if (first >= 0 && k - first < SHK_W) {
            scr_store(A, G, 1 + k - first, true);
            A.virScr[(long)(1 + k - first) * A.nGrp + g] = vir;
        }
        if (cv && prev) {
            settled = true;
            break;
        }
        prev = cv;
        if (k + 1 >= A.P.maxCycle) break;
        shake_r_apply(A, G, dtx, dty, dtz, rt2, vir);
        k++;
    }
    scr_store(A, G, SHK_W + 1, true);
    A.virScr[(long)(SHK_W + 1) * A.nGrp + g] = vir;
    A.first[g] = first;
    A.kend[g] = k;
    kend = k;
}

__global__ void __launch_bounds__(128) k_vv1_a(ShakeArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the wave-wide AND; idle lanes contribute all ones
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) vv1_a_body(A, g, bits, kend, settled);
    shake_publish_mask(A, bits, kend, settled);
}

// Stage 1, part B: every group takes its state at the global stop sweep C
// (from the window, or by sweeping on from the last or the first state).
__global__ void __launch_bounds__(128) k_vv1_b(ShakeArgs A)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nGrp) return;
    if (A.conOff[g + 1] == A.conOff[g]) {
        A.vir[g] = 0.;
        return;
    }
    const int C = shake_global_stop(A);
    if (C < 0) {
        atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 1);
        return;
    }
    Grp G;
    grp_load(A, g, G);
    dd_load(A, G);
    const int first = A.first[g], kend = A.kend[g];
    double vir;
    int at, slot;
    if (first >= 0 && C >= first && C - first < SHK_W && C <= kend) {
        slot = 1 + C - first;
        at = C;
    } else if (C >= kend) {
        slot = SHK_W + 1;
        at = kend;
    } else {
        slot = 0;
        at = 0;
    }
    scr_load(A, G, slot, true);
    vir = slot == 0 ? 0. : A.virScr[(long)slot * A.nGrp + g];
    double dtx[MDG_GRP_CONS], dty[MDG_GRP_CONS], dtz[MDG_GRP_CONS], rt2[MDG_GRP_CONS];
    for (int k = at; k < C; k++) {
        shake_r_check(A, G, dtx, dty, dtz, rt2);
        shake_r_apply(A, G, dtx, dty, dtz, rt2, vir);
    }
    // the state at C passes the check, unless a group re-diverged after
    // it was taken as settled
    if (!(shake_r_check(A, G, dtx, dty, dtz, rt2) < A.P.tolShake))
        atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
    A.vir[g] = vir;
    grp_store(A, G, true);
}

// Stage 2, part A: constraint vectors, half kick, SHAKE-V sweeps with the
// same bookkeeping (integrate.c:1526-1561, shake.c:359-457).
__device__ inline void vv2_a_body(const ShakeArgs &A, int g, unsigned long long (&bits)[SHK_MASKW],
                                  int &kend, bool &settled)
{
#pragma clang fp contract(off)
    Grp G;
    grp_load(A, g, G);
    const double dt = A.P.timeStep;
    if (A.ddStored)
        dd_load(A, G);
    else
        dd_compute(A, G);
    if (!A.noKick)
        for (int k = 0; k < G.na; k++) {
            int i = G.at[k];
            G.vx[k] += 0.5 * dt * A.fx[i] * G.rm[k];
            G.vy[k] += 0.5 * dt * A.fy[i] * G.rm[k];
            G.vz[k] += 0.5 * dt * A.fz[i] * G.rm[k];
        }
    if (G.nc == 0) {
        grp_store(A, G, false);
        for (int w = 0; w < SHK_MASKW; w++) bits[w] = ~0ull;
        kend = -1;
        settled = true;
        return;
    }
    scr_store(A, G, 0, false);
    const double tolvel = A.P.tolShake * A.P.rTimeStep;
    for (int w = 0; w < SHK_MASKW; w++) bits[w] = 0;
    int first = -1, k = 0;
    bool prev = false;
    settled = false;
    for (;;) {
        Grp T = G;
        const bool cv = shake_v_sweep(A, T, false, tolvel, true) < tolvel;   // T = G after a sweep
        if (cv) bits[k >> 6] |= 1ull << (k & 63);
        if (cv && first < 0) first = k;
        if (first >= 0 && k - first < SHK_W) scr_store(A, G, 1 + k - first, false);
        if (cv && prev) {
            settled = true;
            break;
        }
        prev = cv;
        if (k + 1 >= A.P.maxCycle) break;
        G = T;
        k++;
    }
    scr_store(A, G, SHK_W + 1, false);
    A.first[g] = first;
    A.kend[g] = k;
    kend = k;
}

__global__ void __launch_bounds__(128) k_vv2_a(ShakeArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane reaches the wave-wide AND; idle lanes contribute all ones
    unsigned long long bits[SHK_MASKW] = {~0ull, ~0ull, ~0ull, ~0ull};
    int kend = -1;
    bool settled = true;
    if (g < A.nGrp) vv2_a_body(A, g, bits, kend, settled);
    shake_publish_mask(A, bits, kend, settled);
}

// Stage 2, part B: state at the global stop sweep, kinetic energy and
// kinetic stress partials (utils.c:502-542), wrap (integrate.c:1620-1627).
__global__ void __launch_bounds__(128) k_vv2_b(ShakeArgs A)
{
#pragma clang fp contract(off)
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (g < A.nGrp) {
        Grp G;
        grp_load(A, g, G);
        if (A.conOff[g + 1] != A.conOff[g]) {
            const int C = shake_global_stop(A);
            if (C < 0) {
                atomicOr(&A.ctrl[CTRL_SHAKE_FAIL], 2);
            } else {
                dd_load(A, G);
                const int first = A.first[g], kend = A.kend[g];
                int at, slot;
                if (first >= 0 && C >= first && C - first < SHK_W && C <= kend) {
                    slot = 1 + C - first;
                    at = C;
                } else if (C >= kend) {
                    slot = SHK_W + 1;
                    at = kend;
                } else {
                    slot = 0;
                    at = 0;
                }
                scr_load(A, G, slot, false);
                const double tolvel = A.P.tolShake * A.P.rTimeStep;
                for (int k = at; k < C; k++) shake_v_sweep(A, G, false, tolvel, true);
                if (!(shake_v_sweep(A, G, false, tolvel, false) < tolvel))
                    atomicOr(&A.ctrl[CTRL_SHAKE_ORDER], 1);
            }
        }
        for (int k = 0; k < G.na; k++) {
            double m = A.mass[G.at[k]];
            acc[0] += m * (G.vx[k] * G.vx[k] + G.vy[k] * G.vy[k] + G.vz[k] * G.vz[k]);
            acc[1] += m * G.vx[k] * G.vx[k];
            acc[2] += m * G.vx[k] * G.vy[k];
            acc[3] += m * G.vx[k] * G.vz[k];
            acc[4] += m * G.vy[k] * G.vy[k];
            acc[5] += m * G.vy[k] * G.vz[k];
            acc[6] += m * G.vz[k] * G.vz[k];
            // image_update, utils.c:162-218
            double xs = G.x[k], ys = G.y[k], zs = G.z[k];
            if (A.noWrap) {
                // wrapped by a later kernel, with the box the step ends in
            } else if (A.B.type == 1) {
                xs = xs - A.B.a1 * mdg_nint(xs * A.B.u1);
                ys = ys - A.B.a1 * mdg_nint(ys * A.B.u1);
                zs = zs - A.B.a1 * mdg_nint(zs * A.B.u1);
            } else if (A.B.type == 2) {
                xs = xs - A.B.a1 * mdg_nint(xs * A.B.u1);
                ys = ys - A.B.b2 * mdg_nint(ys * A.B.v2);
                zs = zs - A.B.c3 * mdg_nint(zs * A.B.w3);
            } else if (A.B.type == 3) {
                double xt = xs * A.B.u1 + ys * A.B.u2 + zs * A.B.u3;
                double yt = xs * A.B.v1 + ys * A.B.v2 + zs * A.B.v3;
                double zt = xs * A.B.w1 + ys * A.B.w2 + zs * A.B.w3;
                xt -= mdg_nint(xt);
                yt -= mdg_nint(yt);
                zt -= mdg_nint(zt);
                xs = xt * A.B.a1 + yt * A.B.b1 + zt * A.B.c1;
                ys = xt * A.B.a2 + yt * A.B.b2 + zt * A.B.c2;
                zs = xt * A.B.a3 + yt * A.B.b3 + zt * A.B.c3;
            }
            G.x[k] = xs;
            G.y[k] = ys;
            G.z[k] = zs;
        }
        grp_store(A, G, true);
    }
    mdg_block_partials<7>(acc, A.part);
}

__global__ void k_ctrl_reset(int *ctrl, unsigned long long *mask)
{
    if (threadIdx.x < 2 * SHK_MASKW) mask[threadIdx.x] = ~0ull;
}

// Connected components of the constraint graph (host, once).
int mdg_build_groups(mdg_ctx *ctx, const MDG_CONSTRAINT *cons, int nConst)
{
    const int n = ctx->n;
    std::vector<int> parent(n);
    for (int i = 0; i < n; i++) parent[i] = i;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (int c = 0; c < nConst; c++) {
        int a = find(cons[c].a), b = find(cons[c].b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
    std::vector<int> gid(n, -1), root2g(n, -1);
    int ng = 0;
    for (int i = 0; i < n; i++) {
        int r = find(i);
        if (root2g[r] < 0) root2g[r] = ng++;
        gid[i] = root2g[r];
    }
    std::vector<int> aoff(ng + 1, 0), coff(ng + 1, 0);
    for (int i = 0; i < n; i++) aoff[gid[i] + 1]++;
    for (int c = 0; c < nConst; c++) coff[gid[cons[c].a] + 1]++;
    for (int g = 0; g < ng; g++) {
        aoff[g + 1] += aoff[g];
        coff[g + 1] += coff[g];
        if (aoff[g + 1] - aoff[g] > MDG_GRP_ATOMS || coff[g + 1] - coff[g] > MDG_GRP_CONS) {
            ctx->err = "constraint group larger than 8 atoms / 16 constraints";
            return MDG_EUNSUPPORTED;
        }
    }
    std::vector<int> atoms(n), conl(std::max(nConst, 1));
    std::vector<int> fa(aoff.begin(), aoff.end() - 1), fc(coff.begin(), coff.end() - 1);
    for (int i = 0; i < n; i++) atoms[fa[gid[i]]++] = i;       // ascending within a group
    for (int c = 0; c < nConst; c++) conl[fc[gid[cons[c].a]]++] = c;   // global order
    // every group a three-site molecule with constraints (a0,a1), (a0,a2),
    // (a1,a2) in that order: use the register-resident kernels
    bool w3 = ng > 0;
    for (int g = 0; g < ng && w3; g++) {
        if (aoff[g + 1] - aoff[g] != 3 || coff[g + 1] - coff[g] != 3) { w3 = false; break; }
        const int *a = &atoms[aoff[g]];
        const int *c = &conl[coff[g]];
        const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
        for (int k = 0; k < 3; k++)
            if (cons[c[k]].a != a[pa[k]] || cons[c[k]].b != a[pb[k]]) w3 = false;
    }
    ctx->shakeW3 = w3;
    ctx->nGrp = ng;
    if (ctx->shkScr) {
        hipFree(ctx->shkScr); hipFree(ctx->shkVir); hipFree(ctx->shkFirst); hipFree(ctx->shkKend);
        hipFree(ctx->shkRing);
        hipFree(ctx->shkMask);
    }
    HIPCHK(hipMalloc(&ctx->shkScr, (size_t)SHK_SLOTS * 6 * n * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->shkVir, (size_t)SHK_SLOTS * 3 * std::max(ng, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->shkFirst, std::max(ng, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->shkKend, std::max(ng, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->shkRing, SHK_R * std::max(ng, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->shkMask, 2 * SHK_MASKW * sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&ctx->grpAtomOff, (ng + 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->grpConOff, (ng + 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->grpAtoms, n * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->grpCons, std::max(nConst, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->grpCycle, ng * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->grpVir, ng * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->grpAtomOff, aoff.data(), (ng + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->grpConOff, coff.data(), (ng + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->grpAtoms, atoms.data(), n * sizeof(int), hipMemcpyHostToDevice));
    if (nConst > 0)
        HIPCHK(hipMemcpy(ctx->grpCons, conl.data(), nConst * sizeof(int), hipMemcpyHostToDevice));
    return MDG_OK;
}

// Blocks of the VV stage kernels (and so rows of their kinetic partials):
// four lanes per three-site molecule (one lane at SHAKE spread 0), one
// thread per generic group.
int mdg_vv_blocks(const mdg_ctx *ctx)
{
    if (!ctx->shakeW3) return nblk(ctx->nGrp, 128);
    if (ctx->tune.shkSpread == 0) return nblk(ctx->nGrp, 256);
    const long waves = nblk(ctx->nGrp, 16 / ctx->tune.shkSpread);
    return nblk(waves * 64, 256);
}

// Rows of the stage-1 SHAKE virial table: one per block for three-site
// molecules of a single system, one per group otherwise.
int mdg_vir_rows(const mdg_ctx *ctx)
{
    return (ctx->shakeW3 && !ctx->bt) ? mdg_vv_blocks(ctx) : ctx->nGrp;
}

// Runs one VV stage on the device state with forces fx.. (device, atom
// order).  Per-group SHAKE virials go to grpVir (stage 1) and per-block
// kinetic partials to epart + PART_KIN (stage 2); with sync the sums land
// in scal[S_VIRSHAKE] and scal[S_KIN..S_KIN+6] and are returned.
int mdg_vv_stage(mdg_ctx *ctx, int stage, double *fx, double *fy, double *fz, double *kin_out,
                 double *virshake_out, bool sync, const FinArgs *fin, const VvOpts *opt)
{
    if (opt && (!sync || ctx->bt)) {
        ctx->err = "integrator options need a synchronous single-system stage";
        return MDG_EARG;
    }
    ShakeArgs A;
    A.P = ctx->P;
    A.B = ctx->B;
    A.nGrp = ctx->nGrp;
    A.atomOff = ctx->grpAtomOff;
    A.atoms = ctx->grpAtoms;
    A.conOff = ctx->grpConOff;
    A.cons = ctx->grpCons;
    A.ca = ctx->ca;
    A.cb = ctx->cb;
    A.rc2 = ctx->rc2;
    A.thr = ctx->shkThr;
    A.rmass = ctx->rmass;
    A.mass = ctx->mass;
    A.nAtConst = ctx->nAtConst;
    A.x = ctx->x;
    A.y = ctx->y;
    A.z = ctx->z;
    A.vx = ctx->vx;
    A.vy = ctx->vy;
    A.vz = ctx->vz;
    A.fx = fx;
    A.fy = fy;
    A.fz = fz;
    A.dd = ctx->dd;
    A.nConst = ctx->nConst;
    A.cycle = ctx->grpCycle;
    A.vir = ctx->grpVir;
    A.ctrl = ctx->ctrl;
    A.part = ctx->epart + PART_KIN;
    if (!ctx->shkScr) {
        ctx->err = "constraint groups are not set up";
        return MDG_EARG;
    }
    if (ctx->P.maxCycle > 64 * SHK_MASKW) {
        ctx->err = "maxCycle above 256 is not supported";
        return MDG_EUNSUPPORTED;
    }
    A.scr = ctx->shkScr;
    A.virScr = ctx->shkVir;
    A.first = ctx->shkFirst;
    A.kend = ctx->shkKend;
    A.ring = ctx->shkRing;
    // each stage has its own mask (stage 2's beside stage 1's), so a step's
    // record may run while the next step's stage 1 publishes its sweeps
    A.mask = ctx->shkMask + (stage == 1 ? 0 : SHK_MASKW);
    A.boxes = nullptr;
    A.grep = nullptr;
    A.grpKin = nullptr;
    A.spread = ctx->tune.shkSpread;
    A.slotWrite = 0;
    A.track = 0;
    A.atomSlot = nullptr;
    A.sx = A.sy = A.sz = nullptr;
    A.fix = nullptr;
    A.fstr = 0;
    A.dispPart = nullptr;
    A.spr = nullptr;
    A.cO = nullptr;
    A.ncluO = 0;
    A.relPart = nullptr;
    if (opt) {
        A.ddStored = opt->ddStored;
        A.noKick = opt->noKick;
        A.npt = opt->npt;
        A.noWrap = opt->noWrap;
        A.dscale = opt->dscale;
        A.com0 = opt->com[0];
        A.com1 = opt->com[1];
        A.com2 = opt->com[2];
    }
    const bool w3 = ctx->shakeW3 && !opt;             // options run the generic kernels
    const bool m3 = w3 && ctx->tune.shkSpread == 0;   // one molecule per lane
    const bool batch = ctx->bt != nullptr;
    if (stage == 1 && ctx->slotsInStage1 && w3 && !batch && !sync) {
        // a device step without a rebuild: the replay kernel writes the slots
        A.slotWrite = 1;
        A.track = ctx->pruneSkin > 0. && !ctx->pruneFresh ? 1 : 0;
        ctx->pruneFresh = false;
        A.atomSlot = ctx->cl.atomSlot;
        A.sx = ctx->sx;
        A.sy = ctx->sy;
        A.sz = ctx->sz;
        A.fix = ctx->fix;
        A.fstr = ctx->fiy - ctx->fix;
        A.dispPart = ctx->dispPart;
        if (ctx->cl.cO && ctx->spr) {
            A.spr = ctx->spr;
            A.cO = ctx->cl.cO;
            A.ncluO = ctx->cl.nclu;
            A.relPart = ctx->relPart;
        }
        ctx->slotsDone = true;
    }
    if (batch) {
        if (!w3 || sync) {
            ctx->err = "replica batches integrate rigid three-site molecules on the device";
            return MDG_EUNSUPPORTED;
        }
        A.boxes = ctx->bt->boxes;
        A.grep = ctx->bt->grep;
        A.grpKin = ctx->bt->grpKin;
        A.mask = ctx->bt->masks + (stage == 1 ? 0 : (long)ctx->bt->R * SHK_MASKW);
    }
    const int bs = w3 ? 256 : 128;
    const int nb = w3 ? mdg_vv_blocks(ctx) : nblk(ctx->nGrp, 128);
    // The device-resident step resets the mask in the kernels that run
    // before each stage and reduces the partials in its record kernel; the
    // drop-in call does both here.
    if (sync) {
        hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, ctx->stream, ctx->ctrl, ctx->shkMask);
        LAUNCHCHK();
    }
    if (stage == 1) {
        if (w3) {
            // spare blocks: the previous step's record and list prune
            SideWork S{};
            S.nbMain = nb;
            if (ctx->sidePending) {
                S.rebuilt = ctx->sideRebuilt ? 1 : 0;
                if (batch) {
                    S.nbRec = ctx->bt->R;
                    S.BR = mdg_b_rec_args(ctx);
                } else {
                    S.nbRec = 1;
                    S.R = mdg_rec_args(ctx, nb, mdg_vir_rows(ctx), mdg_excl_rows(ctx));
                    S.ctrl = ctx->ctrl;
                    S.out = ctx->recDev;
                }
                if (mdg_prune_on(ctx)) {
                    S.nbPrune = std::min(mdg_prune_blocks(ctx), 2048);
                    S.Pr = mdg_prune_args(ctx);
                }
                ctx->sidePending = false;
            }
            hipLaunchKernelGGL(m3 ? (batch ? k_m3_vv1_a<true> : k_m3_vv1_a<false>)
                                  : (batch ? k_w3_vv1_a<true> : k_w3_vv1_a<false>),
                               dim3(nb + S.nbRec + S.nbPrune), dim3(bs), 0, ctx->stream, A, S);
        } else {
            hipLaunchKernelGGL(k_vv1_a, dim3(nb), dim3(bs), 0, ctx->stream, A);
        }
        LAUNCHCHK();
        // the previous step's record (on stream2) reads the virial rows
        // that the next kernel writes
        if (ctx->recPending) {
            HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evRecDone, 0));
            ctx->recPending = false;
        }
        hipLaunchKernelGGL(m3      ? (batch ? k_m3_vv1_b<true> : k_m3_vv1_b<false>)
                           : batch ? k_w3_vv1_b<true>
                           : w3    ? k_w3_vv1_b<false>
                                   : k_vv1_b,
                           dim3(nb), dim3(bs), 0, ctx->stream, A);
        LAUNCHCHK();
        // SHAKE virial: sum of the per-group (generic) or per-block (three-site)
        // sums, fixed order
        if (sync)
            RETCHK(mdg_reduce_partials(ctx, ctx->grpVir, w3 ? mdg_vir_rows(ctx) : ctx->nGrp, 1,
                                       ctx->scal + S_VIRSHAKE));
    } else {
        if (w3) {
            FinArgs F{};
            int fuse = 0;
            if (fin && !batch && !sync) {
                F = *fin;
                fuse = ctx->precision == MDG_MIXED ? 1 : 2;
            }
            hipLaunchKernelGGL(m3 ? (batch ? k_m3_vv2_a<true> : k_m3_vv2_a<false>)
                                  : (batch ? k_w3_vv2_a<true> : k_w3_vv2_a<false>),
                               dim3(nb), dim3(bs), 0, ctx->stream, A, F, fuse);
        } else {
            hipLaunchKernelGGL(k_vv2_a, dim3(nb), dim3(bs), 0, ctx->stream, A);
        }
        LAUNCHCHK();
        hipLaunchKernelGGL(m3      ? (batch ? k_m3_vv2_b<true> : k_m3_vv2_b<false>)
                           : batch ? k_w3_vv2_b<true>
                           : w3    ? k_w3_vv2_b<false>
                                   : k_vv2_b,
                           dim3(nb), dim3(bs), 0, ctx->stream, A);
        LAUNCHCHK();
        if (sync) RETCHK(mdg_reduce_partials(ctx, A.part, nb, 7, ctx->scal + S_KIN));
    }
    if (sync) {
        int c[CTRL_WORDS];
        double s[8];
        HIPCHK(hipMemcpyAsync(c, ctx->ctrl, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(s, ctx->scal + S_KIN, 8 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        double vs = 0.;
        HIPCHK(hipMemcpyAsync(&vs, ctx->scal + S_VIRSHAKE, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (kin_out) {
            for (int k = 0; k < 7; k++) kin_out[k] = s[k];
            kin_out[0] *= 0.5;
        }
        if (virshake_out) *virshake_out = vs;
        if (c[CTRL_SHAKE_FAIL]) {
            ctx->err = "SHAKE did not converge within maxCycle";
            return 311;
        }
        if (c[CTRL_SHAKE_ORDER]) {
            ctx->err = "a constraint group re-diverged after converging";
            return MDG_ESHAKE_ORDER;
        }
    }
    return MDG_OK;
}
