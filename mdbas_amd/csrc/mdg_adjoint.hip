// mdg_adjoint.hip -- gradients of the potential energy with respect to the
// per-atom nonbonded parameters (charges q, LJ eps and sig as stored by
// io.c:1174-1175), FP64:
//
//  * mdg_backward: dE_pot/d(q, eps, sig) of one configuration (the energy
//    of energy.c:450-643 with spme.c:357-699): real-space pairs over the
//    device cluster list (k_adj_cluster), the SPME mesh potential
//    (k_spme_potential), the excluded-pair correction, the self and
//    net-charge terms (k_adj_finalize).
//
//  * mdg_backward_traj: d/d(q, eps, sig) of the potential energy averaged
//    over steps 1..T of the device-resident VV + SHAKE trajectory
//    (mdbas.c:244-301), by a reverse sweep over the steps.  Positions,
//    velocities and forces are checkpointed every K steps in the forward
//    run; each segment is recomputed forward from its checkpoint, then
//    swept backward.  Per step the adjoint needs
//      - the SHAKE maps' vector-Jacobian products, from the implicit
//        function theorem on each rigid three-site molecule (the constraint
//        solution that the reference's iteration converges to);
//      - (dF/dx)^T w = -H w and (dF/dtheta)^T w, analytic: the pair
//        terms' second radial derivatives over the device list, the
//        excluded-pair correction, and SPME through the mesh of
//        sum_l q_l grad W_l . w_l and the splines' second derivatives, with
//        the terms of the mean-force removal of spme.c:648-670.
#include "mdg_internal.h"

namespace {

constexpr int CA[3] = {0, 0, 1};
constexpr int CB[3] = {1, 2, 2};

struct AdjArgs {
    MdgParams P;
    MdgBox B;
    int n;
    const int *atomSlot;
    const double *gq, *ge, *gs;     // real-space slot sums
    const double *pot;              // dE_rec/dq
    const int *exOff, *exLst;       // symmetric exclusion table
    const double *x, *y, *z, *q;
    double selfc, netc;             // dE_self/dq_i = selfc q_i, dE_net/dq_i = netc
    double scale;
    int accumulate;
    double *oq, *oe, *os;
};

// Per atom: real-space slot sums + reciprocal + excluded-pair correction
// (ewald.c:371-384, -C q_i q_j erf(a r)/r) + self + net charge, scaled.
__global__ void __launch_bounds__(256) k_adj_finalize(AdjArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int s = A.atomSlot[i];
    double gq = A.gq[s] + A.pot[i] + A.selfc * A.q[i] + A.netc;
    for (int k = A.exOff[i]; k < A.exOff[i + 1]; k++) {
        const int j = A.exLst[k];
        const int lo = min(i, j), hi = max(i, j);
        double dx = A.x[hi] - A.x[lo], dy = A.y[hi] - A.y[lo], dz = A.z[hi] - A.z[lo];
        const double r = sqrt(mdg_dist(A.B, dx, dy, dz));
        gq += -A.P.chargeConst * A.q[j] * erf(A.P.alpha * r) / r;
    }
    const double ge = A.ge[s], gs = A.gs[s];
    if (A.accumulate) {
        A.oq[i] += A.scale * gq;
        A.oe[i] += A.scale * ge;
        A.os[i] += A.scale * gs;
    } else {
        A.oq[i] = A.scale * gq;
        A.oe[i] = A.scale * ge;
        A.os[i] = A.scale * gs;
    }
}

// ------------------------------------------------------------ trajectory

// One step's saved state: x, y, z, vx, vy, vz, fx, fy, fz (n each).
struct StepState {
    const double *x[3], *v[3], *f[3];
};

__device__ inline StepState step_state(const double *base, int n)
{
    StepState S;
    for (int c = 0; c < 3; c++) {
        S.x[c] = base + (long)c * n;
        S.v[c] = base + (long)(3 + c) * n;
        S.f[c] = base + (long)(6 + c) * n;
    }
    return S;
}

struct TrajArgs {
    MdgBox B;
    int nGrp, n;
    double dt;
    const int *atomOff, *atoms;
    const double *rmass;
    const double *prev, *cur;       // saved states t-1 and t
    double *xb;                     // adjoint of x_t (3n), updated to that of x_{t-1}
    double *vb;                     // adjoint of v_t (3n), updated to that of v_{t-1}
    double *vb2;                    // adjoint of v'' (velocity after SHAKE-R), 3n
    double *fp;                     // adjoint of F_{t-1} from stage 1 (3n)
};

// Minimum image of a short vector (utils.c:38-80).
__device__ inline void minimg(const MdgBox &B, double d[3])
{
    mdg_dist(B, d[0], d[1], d[2]);
}

__device__ inline double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Solves M y = r for a 3x3 matrix by cofactors; trans solves M^T y = r.
__device__ inline void solve3(const double M[3][3], const double r[3], double y[3], bool trans)
{
    double A[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A[i][j] = trans ? M[j][i] : M[i][j];
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    const double inv[3][3] = {
        {c00, A[0][2] * A[2][1] - A[0][1] * A[2][2], A[0][1] * A[1][2] - A[0][2] * A[1][1]},
        {c01, A[0][0] * A[2][2] - A[0][2] * A[2][0], A[0][2] * A[1][0] - A[0][0] * A[1][2]},
        {c02, A[0][1] * A[2][0] - A[0][0] * A[2][1], A[0][0] * A[1][1] - A[0][1] * A[1][0]}};
    for (int i = 0; i < 3; i++)
        y[i] = (inv[i][0] * r[0] + inv[i][1] * r[1] + inv[i][2] * r[2]) / det;
}

// Molecule g's state at t-1 and t, and the quantities of the forward step
// (integrate.c:1518-1634): v' = v_{t-1} + dt/2 F_{t-1}/m, x~ = x_{t-1} +
// dt v', dx = x_t - x~ (the SHAKE-R displacement), v'' = v' + dx/dt,
// v~ = v'' + dt/2 F_t/m (before SHAKE-V).
struct Mol {
    int at[3];
    double rm[3];
    double xp[3][3], xt[3][3], vt[3][3], dx[3][3], vtil[3][3];
};

__device__ inline void mol_load(const TrajArgs &A, int g, Mol &M)
{
    const StepState P = step_state(A.prev, A.n), C = step_state(A.cur, A.n);
    const int o = A.atomOff[g];
    for (int k = 0; k < 3; k++) {
        const int i = A.atoms[o + k];
        M.at[k] = i;
        M.rm[k] = A.rmass[i];
        double d[3];
        for (int c = 0; c < 3; c++) {
            const double vp = P.v[c][i] + 0.5 * A.dt * P.f[c][i] * M.rm[k];
            const double xtil = P.x[c][i] + A.dt * vp;
            M.xp[k][c] = P.x[c][i];
            M.xt[k][c] = C.x[c][i];
            M.vt[k][c] = C.v[c][i];
            d[c] = C.x[c][i] - xtil;   // the wrap of stage 2 is a lattice shift
            M.vtil[k][c] = vp;         // completed below
        }
        minimg(A.B, d);
        for (int c = 0; c < 3; c++) {
            M.dx[k][c] = d[c];
            M.vtil[k][c] += d[c] / A.dt + 0.5 * A.dt * C.f[c][i] * M.rm[k];
        }
    }
}

// Constraint direction rows: r_c = x_b - x_a (minimum image); row c of the
// Jacobian is -r_c on atom a and +r_c on atom b.
__device__ inline void mol_dirs(const MdgBox &B, const double X[3][3], double r[3][3])
{
    for (int c = 0; c < 3; c++) {
        for (int k = 0; k < 3; k++) r[c][k] = X[CB[c]][k] - X[CA[c]][k];
        minimg(B, r[c]);
    }
}

// (R M^-1 D^T)[c][d] for rows r (of R) and d (of D).
__device__ inline void mol_gram(const double r[3][3], const double d[3][3], const double rm[3],
                                double G[3][3])
{
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 3; e++) {
            double s = 0.;
            for (int k = 0; k < 3; k++) {
                // atom k's coefficient in row c of R and row e of D
                const double sr = (k == CA[c]) ? -1. : (k == CB[c]) ? 1. : 0.;
                const double sd = (k == CA[e]) ? -1. : (k == CB[e]) ? 1. : 0.;
                if (sr != 0. && sd != 0.) s += sr * sd * rm[k] * dot3(r[c], d[e]);
            }
            G[c][e] = s;
        }
}

// (R u)_c = r_c . (u_b - u_a) with optional per-atom weights w.
__device__ inline void mol_apply(const double r[3][3], const double u[3][3], const double *w,
                                 double out[3])
{
    for (int c = 0; c < 3; c++) {
        double d[3];
        const double wa = w ? w[CA[c]] : 1., wb = w ? w[CB[c]] : 1.;
        for (int k = 0; k < 3; k++) d[k] = wb * u[CB[c]][k] - wa * u[CA[c]][k];
        out[c] = dot3(r[c], d);
    }
}

// Adds R^T z to u (per atom: -z_c r_c on a, +z_c r_c on b), times s.
__device__ inline void mol_add_rt(const double r[3][3], const double z[3], double s, double u[3][3])
{
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            u[CA[c]][k] -= s * z[c] * r[c][k];
            u[CB[c]][k] += s * z[c] * r[c][k];
        }
}

}  // namespace

// Stage 2 (SHAKE-V, shake.c:359-457, as the projection v = v~ + M^-1 B^T mu
// with B(x_t) v = 0): from the adjoint of v_t, the adjoint of v~ (to vb2)
// and the x_t part through B (added to xb).
__global__ void __launch_bounds__(64) k_traj_stage2(TrajArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nGrp) return;
    Mol M;
    mol_load(A, g, M);
    double r[3][3], S[3][3];
    mol_dirs(A.B, M.xt, r);
    mol_gram(r, r, M.rm, S);
    double dv[3][3], vbar[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            dv[k][c] = M.vt[k][c] - M.vtil[k][c];
            vbar[k][c] = A.vb[(long)c * A.n + M.at[k]];
        }
    double t[3], mu[3], kap[3];
    mol_apply(r, dv, nullptr, t);
    solve3(S, t, mu, false);
    mol_apply(r, vbar, M.rm, t);
    solve3(S, t, kap, false);
    // adjoint of v~: v_bar - B^T kappa
    double vt_bar[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) vt_bar[k][c] = vbar[k][c];
    mol_add_rt(r, kap, -1., vt_bar);
    // x_t through B: G_c = mu_c (w_b - w_a) - kappa_c (v_b - v_a), w = M^-1 vt_bar
    for (int c = 0; c < 3; c++) {
        double G[3];
        for (int k = 0; k < 3; k++)
            G[k] = mu[c] * (M.rm[CB[c]] * vt_bar[CB[c]][k] - M.rm[CA[c]] * vt_bar[CA[c]][k]) -
                   kap[c] * (M.vt[CB[c]][k] - M.vt[CA[c]][k]);
        for (int k = 0; k < 3; k++) {
            A.xb[(long)k * A.n + M.at[CB[c]]] += G[k];
            A.xb[(long)k * A.n + M.at[CA[c]]] -= G[k];
        }
    }
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) A.vb2[(long)c * A.n + M.at[k]] = vt_bar[k][c];
}

// Stage 1 (drift + SHAKE-R, shake.c:209-357, as x_t = x~ + M^-1 A(x_{t-1})^T
// gamma with the constraints met at x_t, and v'' = v' + (x_t - x~)/dt):
// from the adjoints of x_t (xb) and v'' (vb2), the adjoints of x_{t-1}
// (to xb), v_{t-1} (to vb) and F_{t-1} (to fp).
__global__ void __launch_bounds__(64) k_traj_stage1(TrajArgs A)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.nGrp) return;
    Mol M;
    mol_load(A, g, M);
    double dd[3][3], r[3][3], Q[3][3], T[3][3];
    mol_dirs(A.B, M.xp, dd);
    mol_dirs(A.B, M.xt, r);
    mol_gram(dd, dd, M.rm, Q);
    mol_gram(r, dd, M.rm, T);
    double t[3], gam[3], z[3];
    mol_apply(dd, M.dx, nullptr, t);
    solve3(Q, t, gam, false);
    double xe[3][3], v2b[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            v2b[k][c] = A.vb2[(long)c * A.n + M.at[k]];
            xe[k][c] = A.xb[(long)c * A.n + M.at[k]] + v2b[k][c] / A.dt;
        }
    // u_bar = Pi^T xe = xe - R^T T^-T (A M^-1 xe)
    mol_apply(dd, xe, M.rm, t);
    solve3(T, t, z, true);
    double ub[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) ub[k][c] = xe[k][c];
    mol_add_rt(r, z, -1., ub);
    // adjoint of x~ and of x_{t-1}
    double xt_bar[3][3], xp_bar[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            xt_bar[k][c] = ub[k][c] - v2b[k][c] / A.dt;
            xp_bar[k][c] = xt_bar[k][c];
        }
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            const double coef = gam[c] * (M.rm[CB[c]] * ub[CB[c]][k] - M.rm[CA[c]] * ub[CA[c]][k]);
            xp_bar[CB[c]][k] += coef;
            xp_bar[CA[c]][k] -= coef;
        }
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) {
            const long o = (long)c * A.n + M.at[k];
            const double vpb = v2b[k][c] + A.dt * xt_bar[k][c];
            A.xb[o] = xp_bar[k][c];
            A.vb[o] = vpb;
            A.fp[o] = 0.5 * A.dt * M.rm[k] * vpb;
        }
}

// y = a + s b over m elements.
__global__ void k_axpy(long m, const double *a, double s, const double *b, double *y)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) y[i] = a[i] + s * b[i];
}

// fbar = dt/2 vb2 / m + fp (per component).
__global__ void __launch_bounds__(256) k_fbar(int n, double dt, const double *rmass,
                                              const double *vb2, const double *fp, double *fbar)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; c++) {
        const long o = (long)c * n + i;
        fbar[o] = 0.5 * dt * rmass[i] * vb2[o] + fp[o];
    }
}

// ------------------------------------------------------------ host side

static int ensure_adj(mdg_ctx *ctx)
{
    if (ctx->adjQ) return MDG_OK;
    const size_t n = ctx->n;
    HIPCHK(hipMalloc(&ctx->adjQ, n * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->adjE, n * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->adjS, n * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->adjPot, n * sizeof(double)));
    return MDG_OK;
}

// scale * dE/d(q, eps, sig) at the context's current positions into
// adjQ/adjE/adjS (added when accumulate).
static int adj_grad(mdg_ctx *ctx, double scale, bool accumulate)
{
    RETCHK(ensure_adj(ctx));
    RETCHK(mdg_adj_pair(ctx));
    RETCHK(mdg_spme_pot(ctx, ctx->adjPot));
    AdjArgs A;
    A.P = ctx->P;
    A.B = ctx->B;
    A.n = ctx->n;
    A.atomSlot = ctx->cl.atomSlot;
    A.gq = ctx->fdx;
    A.ge = ctx->fdy;
    A.gs = ctx->fdz;
    A.pot = ctx->adjPot;
    A.exOff = ctx->exSymOff;
    A.exLst = ctx->exSymLst;
    A.x = ctx->x;
    A.y = ctx->y;
    A.z = ctx->z;
    A.q = ctx->q;
    A.selfc = -2.0 * ctx->P.chargeConst * ctx->P.alpha / SQRTPI;
    A.netc = ctx->P.alpha > 0 && ctx->B.vol > 0
        ? -(3.141592653589793 * ctx->P.chargeConst) * ctx->systq / (ctx->P.alpha * ctx->P.alpha * ctx->B.vol)
        : 0.;
    A.scale = scale;
    A.accumulate = accumulate ? 1 : 0;
    A.oq = ctx->adjQ;
    A.oe = ctx->adjE;
    A.os = ctx->adjS;
    hipLaunchKernelGGL(k_adj_finalize, dim3(nblk(ctx->n, 256)), dim3(256), 0, ctx->stream, A);
    LAUNCHCHK();
    return MDG_OK;
}

static int adj_download(mdg_ctx *ctx, double dq[], double deps[], double dsig[])
{
    const size_t b = ctx->n * sizeof(double);
    if (dq) HIPCHK(hipMemcpyAsync(dq, ctx->adjQ, b, hipMemcpyDeviceToHost, ctx->stream));
    if (deps) HIPCHK(hipMemcpyAsync(deps, ctx->adjE, b, hipMemcpyDeviceToHost, ctx->stream));
    if (dsig) HIPCHK(hipMemcpyAsync(dsig, ctx->adjS, b, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDG_OK;
}

static int adj_ready(mdg_ctx *ctx)
{
    if (!ctx->cl.cpJ || !ctx->sx || !ctx->exSymOff) {
        ctx->err = "call mdg_initial_energy first (the adjoint uses the device list)";
        return MDG_EARG;
    }
    if (ctx->P.ewald != 2) {
        ctx->err = "the adjoint differentiates the Ewald/SPME potential (keyEwald 2)";
        return MDG_EUNSUPPORTED;
    }
    return MDG_OK;
}

extern "C" int mdg_backward(mdg_ctx *ctx, double dL_dE, double dL_dq[], double dL_deps[],
                            double dL_dsig[])
{
    if (!ctx) return MDG_EARG;
    RETCHK(adj_ready(ctx));
    RETCHK(adj_grad(ctx, dL_dE, false));
    return adj_download(ctx, dL_dq, dL_deps, dL_dsig);
}

static int d2d(mdg_ctx *ctx, double *dst, const double *src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return MDG_OK;
}

// state <-> 9n buffer (x, y, z, vx, vy, vz, fx, fy, fz)
static int save_state(mdg_ctx *ctx, double *buf)
{
    const size_t n = ctx->n;
    double *a[9] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz, ctx->fx, ctx->fy, ctx->fz};
    for (int k = 0; k < 9; k++) RETCHK(d2d(ctx, buf + k * n, a[k], n));
    return MDG_OK;
}

static int load_state(mdg_ctx *ctx, const double *buf)
{
    const size_t n = ctx->n;
    double *a[9] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz, ctx->fx, ctx->fy, ctx->fz};
    for (int k = 0; k < 9; k++) RETCHK(d2d(ctx, a[k], buf + k * n, n));
    return MDG_OK;
}

// ctx positions = base positions (3n buffer) + s u.
static int set_positions(mdg_ctx *ctx, const double *base, double s, const double *u)
{
    const int n = ctx->n;
    double *a[3] = {ctx->x, ctx->y, ctx->z};
    for (int c = 0; c < 3; c++) {
        if (u)
            hipLaunchKernelGGL(k_axpy, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, (long)n,
                               base + (long)c * n, s, u + (long)c * n, a[c]);
        else
            RETCHK(d2d(ctx, a[c], base + (long)c * n, n));
    }
    LAUNCHCHK();
    return MDG_OK;
}

// y[c*n + i] += s * (F_c[i] + m[c])
__global__ void k_add_raw(int n, double s, const double *fx, const double *fy, const double *fz,
                          const double *m, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] += s * (fx[i] + m[0]);
    y[n + i] += s * (fy[i] + m[1]);
    y[2 * n + i] += s * (fz[i] + m[2]);
}

// out[0..2] = means of a, b, c (one block, fixed order).
__global__ void __launch_bounds__(1024) k_mean3(int n, const double *a, const double *b,
                                                const double *c, double *out)
{
    __shared__ double red[16][3];
    double v[3] = {0., 0., 0.};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        v[0] += a[i];
        v[1] += b[i];
        v[2] += c[i];
    }
    for (int k = 0; k < 3; k++) {
        const double w = mdg_wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = w;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w][threadIdx.x];
        out[threadIdx.x] = t / (double)n;
    }
}

// u in slot order: su[c * ns + s] = u[c * n + atom of slot s], 0 on
// padding slots.
__global__ void k_hvp_slot_u(int nslot, long ns, const int *__restrict__ slotAtom, int n,
                             const double *__restrict__ u, double *__restrict__ su)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int a = slotAtom[s];
    for (int c = 0; c < 3; c++) su[c * ns + s] = a >= 0 ? u[(long)c * n + a] : 0.;
}

struct HvpFin {
    MdgParams P;
    MdgBox B;
    int n;
    long ns;
    const int *atomSlot, *exOff, *exLst;   // symmetric exclusion table
    const double *x, *y, *z, *q, *u;
    const double *slot;                    // slot sums: hx, hy, hz, mq, me, ms (ns each)
    double *hu, *mq, *me, *ms;             // atom order (overwritten)
};

// Per atom: the real-space slot sums plus the excluded-pair correction
// (phi = -C q_i q_j erf(a r)/r, ewald.c:371-384) over the symmetric table,
// each partner contributing -(g du + h (d . du) d) to (H u)_i and
// C q_j phi'/(qel r) (d . du) to d/dq_i (grad E . u).
__global__ void __launch_bounds__(256) k_hvp_finalize(HvpFin F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n) return;
    const int s = F.atomSlot[i];
    double h[3] = {F.slot[s], F.slot[F.ns + s], F.slot[2 * F.ns + s]};
    double gq = F.slot[3 * F.ns + s];
    const double C = F.P.chargeConst, al = F.P.alpha;
    const double ui[3] = {F.u[i], F.u[F.n + i], F.u[2 * (long)F.n + i]};
    for (int k = F.exOff[i]; k < F.exOff[i + 1]; k++) {
        const int j = F.exLst[k];
        double d[3] = {F.x[j] - F.x[i], F.y[j] - F.y[i], F.z[j] - F.z[i]};
        const double r = sqrt(mdg_dist(F.B, d[0], d[1], d[2])), rt = 1. / r, ar = al * r;
        const double ef = erf(ar), gs = 2. * al * exp(-ar * ar) / SQRTPI;
        const double c1 = (ef * rt - gs) * rt;                                 // per unit qel
        const double c2 = -2. * ef * rt * rt * rt + gs * (2. * rt * rt + 2. * al * al);
        const double qel = C * F.q[i] * F.q[j];
        const double g = qel * c1 * rt, hh = (qel * c2 - g) * rt * rt;
        const double du[3] = {F.u[j] - ui[0], F.u[F.n + j] - ui[1], F.u[2 * (long)F.n + j] - ui[2]};
        const double t = d[0] * du[0] + d[1] * du[1] + d[2] * du[2];
        for (int c = 0; c < 3; c++) h[c] -= g * du[c] + hh * t * d[c];
        gq += C * F.q[j] * c1 * rt * t;
    }
    for (int c = 0; c < 3; c++) F.hu[(long)c * F.n + i] = h[c];
    F.mq[i] = gq;
    F.me[i] = F.slot[4 * F.ns + s];
    F.ms[i] = F.slot[5 * F.ns + s];
}

// The second-derivative scratch: u in slot order (3 ns), the slot sums
// (6 ns), then in atom order H u (3n), its mixed terms (3n) and the
// reciprocal part along the uniform translation (3n + n).
static long hvp_ns(const mdg_ctx *ctx) { return (long)ctx->cl.nclu * MDG_CL + 64; }

static int ensure_hvp(mdg_ctx *ctx)
{
    if (ctx->hvpBuf) return MDG_OK;
    const long ns = hvp_ns(ctx), n = ctx->n;
    HIPCHK(hipMalloc(&ctx->hvpBuf, (size_t)(9 * ns + 10 * n) * sizeof(double)));
    return MDG_OK;
}

// Scratch of the force vector-Jacobian product.
struct VjpScratch {
    double *m;       // 8: [0..2] mean raw reciprocal force, [4..6] mean of u
};

// Adds (dF/dx)^T u to xb (when xb) and (dF/dtheta)^T u to the parameter
// adjoints at positions base (3n).  F = Fr - m with Fr = -grad E and m the
// mean raw reciprocal force that spme.c:648-670 removes, so
//   (dF/dx)^T u     = -H u + (1/N) H_rec (s, ..., s),          s = sum_i u_i
//   (dF/dtheta)^T u = -d/dtheta (grad E . u) + (1/N) d/dq (grad E_rec . (s, ..., s))
// with H u and the mixed terms analytic: the real-space pairs over the
// device list (k_hvp_cluster), the excluded-pair correction
// (k_hvp_finalize) and SPME (mdg_spme_hvp: the dipole mesh of u through
// the influence function, and the splines' second derivatives).
static int force_vjp(mdg_ctx *ctx, const double *base, const double *u, double *xb,
                     const VjpScratch &W)
{
    const int n = ctx->n, nbA = nblk(n, 256);
    const long ns = hvp_ns(ctx);
    RETCHK(ensure_hvp(ctx));
    RETCHK(set_positions(ctx, base, 0., nullptr));
    double *su = ctx->hvpBuf, *slot = su + 3 * ns;
    double *hu = slot + 6 * ns, *mq = hu + 3 * (long)n, *me = mq + n, *ms = me + n;
    double *hu2 = ms + n, *mq2 = hu2 + 3 * (long)n;
    RETCHK(mdg_fill_f64(ctx, slot, 6 * ns, 0.0));
    const int nslot = ctx->cl.nclu * MDG_CL;
    hipLaunchKernelGGL(k_hvp_slot_u, dim3(nblk(nslot, 256)), dim3(256), 0, ctx->stream, nslot, ns,
                       ctx->cl.slotAtom, n, u, su);
    LAUNCHCHK();
    HvpSlots H;
    H.ux = su;
    H.uy = su + ns;
    H.uz = su + 2 * ns;
    H.hx = slot;
    H.hy = slot + ns;
    H.hz = slot + 2 * ns;
    H.mq = slot + 3 * ns;
    H.me = slot + 4 * ns;
    H.ms = slot + 5 * ns;
    RETCHK(mdg_hvp_pair(ctx, H));
    HvpFin F;
    F.P = ctx->P;
    F.B = ctx->B;
    F.n = n;
    F.ns = ns;
    F.atomSlot = ctx->cl.atomSlot;
    F.exOff = ctx->exSymOff;
    F.exLst = ctx->exSymLst;
    F.x = ctx->x;
    F.y = ctx->y;
    F.z = ctx->z;
    F.q = ctx->q;
    F.u = u;
    F.slot = slot;
    F.hu = hu;
    F.mq = mq;
    F.me = me;
    F.ms = ms;
    hipLaunchKernelGGL(k_hvp_finalize, dim3(nbA), dim3(256), 0, ctx->stream, F);
    LAUNCHCHK();
    RETCHK(mdg_spme_hvp(ctx, u, nullptr, hu, mq));
    // -H u and -d/dtheta (grad E . u)
    if (xb)
        hipLaunchKernelGGL(k_axpy, dim3(nblk(3L * n, 256)), dim3(256), 0, ctx->stream, 3L * n, xb,
                           -1.0, hu, xb);
    hipLaunchKernelGGL(k_axpy, dim3(nbA), dim3(256), 0, ctx->stream, (long)n, ctx->adjQ, -1.0, mq,
                       ctx->adjQ);
    hipLaunchKernelGGL(k_axpy, dim3(nbA), dim3(256), 0, ctx->stream, (long)n, ctx->adjE, -1.0, me,
                       ctx->adjE);
    hipLaunchKernelGGL(k_axpy, dim3(nbA), dim3(256), 0, ctx->stream, (long)n, ctx->adjS, -1.0, ms,
                       ctx->adjS);
    // the mean removal's terms along the uniform translation s
    hipLaunchKernelGGL(k_mean3, dim3(1), dim3(1024), 0, ctx->stream, n, u, u + n, u + 2 * (long)n,
                       W.m + 4);
    LAUNCHCHK();
    double sm[3];
    HIPCHK(hipMemcpyAsync(sm, W.m + 4, sizeof sm, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double sv[3] = {sm[0] * n, sm[1] * n, sm[2] * n};
    RETCHK(mdg_fill_f64(ctx, hu2, 4L * n, 0.0));
    RETCHK(mdg_spme_hvp(ctx, nullptr, sv, hu2, mq2));
    const double inv = 1.0 / n;
    if (xb)
        hipLaunchKernelGGL(k_axpy, dim3(nblk(3L * n, 256)), dim3(256), 0, ctx->stream, 3L * n, xb,
                           inv, hu2, xb);
    hipLaunchKernelGGL(k_axpy, dim3(nbA), dim3(256), 0, ctx->stream, (long)n, ctx->adjQ, inv, mq2,
                       ctx->adjQ);
    LAUNCHCHK();
    return MDG_OK;
}

extern "C" int mdg_backward_traj(mdg_ctx *ctx, int nsteps, int ckpt, double *J, double dJ_dq[],
                                 double dJ_deps[], double dJ_dsig[])
{
    if (!ctx || nsteps <= 0 || ckpt <= 0) return MDG_EARG;
    RETCHK(adj_ready(ctx));
    if (ctx->ens != 0) {
        ctx->err = "the trajectory adjoint differentiates NVE dynamics";
        return MDG_EUNSUPPORTED;
    }
    if (ctx->precision != MDG_FP64) {
        ctx->err = "the trajectory adjoint runs in FP64 contexts";
        return MDG_EUNSUPPORTED;
    }
    if (!ctx->shakeW3 || ctx->hasFrozen) {
        ctx->err = "the trajectory adjoint covers rigid three-site molecules without frozen atoms";
        return MDG_EUNSUPPORTED;
    }
    RETCHK(ensure_adj(ctx));
    const int n = ctx->n, T = nsteps, K = ckpt;
    const int nseg = (T + K - 1) / K;
    const size_t sn = (size_t)9 * n;
    std::vector<double *> bufs;
    auto alloc = [&](size_t cnt) -> double * {
        double *p = nullptr;
        if (hipMalloc(&p, cnt * sizeof(double)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    double *ck = alloc(sn * nseg), *seg = alloc(sn * (K + 1)), *fin = alloc(sn);
    double *xb = alloc(3 * (size_t)n), *vb = alloc(3 * (size_t)n), *vb2 = alloc(3 * (size_t)n);
    double *fp = alloc(3 * (size_t)n), *fbar = alloc(3 * (size_t)n);
    VjpScratch W;
    W.m = alloc(8);
    int st = MDG_OK;
    auto fail = [&](int code) {
        for (double *p : bufs) hipFree(p);
        return code;
    };
    for (double *p : bufs)
        if (!p) {
            ctx->err = "hipMalloc failed (trajectory adjoint buffers)";
            return fail(MDG_EHIP);
        }
    if (bufs.size() != 9) {
        ctx->err = "hipMalloc failed (trajectory adjoint buffers)";
        return fail(MDG_EHIP);
    }
    const int step0 = ctx->step;
    // forward: checkpoints and the objective
    double Jsum = 0.;
    std::vector<MDG_STEP> rec(K);
    for (int c = 0; c < nseg && st == MDG_OK; c++) {
        st = save_state(ctx, ck + sn * c);
        const int k = std::min(K, T - c * K);
        if (st == MDG_OK) st = mdg_run_steps(ctx, k, 0, rec.data());
        for (int s = 0; s < k && st == MDG_OK; s++) Jsum += rec[s].pot;
    }
    if (st == MDG_OK) st = save_state(ctx, fin);
    if (st) return fail(st);
    const double invT = 1.0 / T;
    if (J) *J = Jsum * invT;
    // reverse sweep
    RETCHK(mdg_fill_f64(ctx, ctx->adjQ, n, 0.0));
    RETCHK(mdg_fill_f64(ctx, ctx->adjE, n, 0.0));
    RETCHK(mdg_fill_f64(ctx, ctx->adjS, n, 0.0));
    RETCHK(mdg_fill_f64(ctx, xb, 3 * (size_t)n, 0.0));
    RETCHK(mdg_fill_f64(ctx, vb, 3 * (size_t)n, 0.0));
    RETCHK(mdg_fill_f64(ctx, fp, 3 * (size_t)n, 0.0));
    TrajArgs A;
    A.B = ctx->B;
    A.nGrp = ctx->nGrp;
    A.n = n;
    A.dt = ctx->P.timeStep;
    A.atomOff = ctx->grpAtomOff;
    A.atoms = ctx->grpAtoms;
    A.rmass = ctx->rmass;
    A.xb = xb;
    A.vb = vb;
    A.vb2 = vb2;
    A.fp = fp;
    const int nbA = nblk(n, 256), nbG = nblk(ctx->nGrp, 64);
    for (int c = nseg - 1; c >= 0 && st == MDG_OK; c--) {
        const int k0 = c * K, kn = std::min(K, T - k0);
        // recompute the segment from its checkpoint
        if ((st = load_state(ctx, ck + sn * c))) break;
        ctx->step = step0 + k0;
        if ((st = mdg_device_energy(ctx, true))) break;
        if ((st = save_state(ctx, seg))) break;
        if ((st = mdg_reset_step_state(ctx))) break;
        for (int k = 1; k <= kn && st == MDG_OK; k++) {
            const int step = ctx->step + 1;
            st = mdg_device_step(ctx, (step % ctx->P.update) == 0);
            if (st == MDG_OK) st = save_state(ctx, seg + sn * k);
            ctx->step = step;
        }
        if (st == MDG_OK) st = mdg_check_flags(ctx);
        for (int k = kn; k >= 1 && st == MDG_OK; k--) {
            const double *cur = seg + sn * k, *prev = seg + sn * (k - 1);
            // objective: J = (1/T) sum_t E(x_t); dE/dx = -(F + m), m the
            // mean raw reciprocal force at x_t (left in frx64 by adj_grad)
            if ((st = set_positions(ctx, cur, 0., nullptr))) break;
            if ((st = adj_grad(ctx, invT, true))) break;
            hipLaunchKernelGGL(k_mean3, dim3(1), dim3(1024), 0, ctx->stream, n, ctx->frx64,
                               ctx->fry64, ctx->frz64, W.m);
            hipLaunchKernelGGL(k_add_raw, dim3(nbA), dim3(256), 0, ctx->stream, n, -invT,
                               cur + (size_t)6 * n, cur + (size_t)7 * n, cur + (size_t)8 * n, W.m,
                               xb);
            A.prev = prev;
            A.cur = cur;
            hipLaunchKernelGGL(k_traj_stage2, dim3(nbG), dim3(64), 0, ctx->stream, A);
            hipLaunchKernelGGL(k_fbar, dim3(nbA), dim3(256), 0, ctx->stream, n, A.dt, ctx->rmass,
                               vb2, fp, fbar);
            if ((st = force_vjp(ctx, cur, fbar, xb, W))) break;
            hipLaunchKernelGGL(k_traj_stage1, dim3(nbG), dim3(64), 0, ctx->stream, A);
            if (hipGetLastError() != hipSuccess) {
                ctx->err = "kernel launch failed (trajectory adjoint)";
                st = MDG_EHIP;
            }
        }
        // F_0 of the whole trajectory depends on the parameters too
        if (st == MDG_OK && c == 0) {
            st = force_vjp(ctx, seg, fp, nullptr, W);
        }
    }
    // leave the context at the end of the trajectory
    if (st == MDG_OK) st = load_state(ctx, fin);
    if (st == MDG_OK) {
        ctx->step = step0 + T;
        st = mdg_device_energy(ctx, true);
    }
    if (st == MDG_OK) st = adj_download(ctx, dJ_dq, dJ_deps, dJ_dsig);
    return fail(st);
}
