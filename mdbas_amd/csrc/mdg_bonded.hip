// mdg_bonded.hip -- bonded terms and 1-4 pairs on the device (FP64).
//
//  * bond / Urey-Bradley (internal.c:21-150), angle (164-245), dihedral
//    and improper (247-604), and the 1-4 pairs of nonbond14_energy
//    (energy.c:356-439, coulomb14_* elec.c:272-465, vdw14_switch vdw.c:
//    270-311) or ewald14_energy (energy.c:654-742, ewald_dir14 /
//    ewald_corr14 ewald.c:386-413), in one kernel: one thread per term over
//    the concatenated term ranges, FP64 arithmetic in the reference's
//    expression order, forces added into the per-atom FP64 arrays (atomics)
//    or, for the device-resident mixed step, into the int64 fixed-point
//    per-atom accumulator (integer atomics: order-independent, so the step
//    stays bitwise reproducible).  Energies and virials are per-block
//    partials reduced in a fixed order.
#include "mdg_internal.h"

namespace {

struct BondedArgs {
    MdgParams P;
    MdgBox B;
    const double *x, *y, *z, *q, *eps14, *sig14;
    const MDG_BOND *bond, *ub;
    const MDG_ANGLE *angle;
    const MDG_DIHE *dihe, *impr;
    const int *l14;
    int nBond, nAngle, nUb, nDihe, nImpr, n14;
    double scal14;
    int ewald14, elec14, vdw14;
    double *fx, *fy, *fz;        // FP64 targets (fix == nullptr)
    long long *fix;              // int64 targets fix[0..3n) (x | y | z), MDG_FSCALE
    int n;
    double *part;                // [block][BD_NV]
};

__device__ inline void add_force(const BondedArgs &A, int i, double fx, double fy, double fz)
{
    if (A.fix) {
        mdg_atomic_add_fix(A.fix + i, fx, MDG_FSCALE);
        mdg_atomic_add_fix(A.fix + A.n + i, fy, MDG_FSCALE);
        mdg_atomic_add_fix(A.fix + 2 * (long)A.n + i, fz, MDG_FSCALE);
    } else {
        atomicAdd(&A.fx[i], fx);
        atomicAdd(&A.fy[i], fy);
        atomicAdd(&A.fz[i], fz);
    }
}

// internal.c:33-80 (bond; Morse for type 1) and 120-150 (Urey-Bradley,
// harmonic only): returns the energy, vir += dE/dr * r.
__device__ double bond_term(const BondedArgs &A, const MDG_BOND &t, bool morse, double &vir)
{
#pragma clang fp contract(off)
    const int i = t.a, j = t.b;
    double dx = A.x[j] - A.x[i], dy = A.y[j] - A.y[i], dz = A.z[j] - A.z[i];
    const double r = sqrt(mdg_dist(A.B, dx, dy, dz));
    double e, db;
    if (morse && t.type == 1) {
        const double ma = exp(-t.beta * (r - t.r0));
        const double mb = ma * ma;
        e = t.k * (mb - 2. * ma) + t.k;
        db = 2. * t.k * t.beta * (ma - mb);
    } else {
        const double dr = r - t.r0;
        e = 0.5 * t.k * (dr * dr);
        db = t.k * dr;
    }
    vir += db * r;
    const double tfx = db * dx / r, tfy = db * dy / r, tfz = db * dz / r;
    add_force(A, i, tfx, tfy, tfz);
    add_force(A, j, -tfx, -tfy, -tfz);
    return e;
}

// internal.c:164-245
__device__ double angle_term(const BondedArgs &A, const MDG_ANGLE &t)
{
#pragma clang fp contract(off)
    const int i = t.a, j = t.b, k = t.c;
    double ax = A.x[i] - A.x[j], ay = A.y[i] - A.y[j], az = A.z[i] - A.z[j];
    const double rab = sqrt(mdg_dist(A.B, ax, ay, az)), rabt = 1. / rab;
    double cx = A.x[k] - A.x[j], cy = A.y[k] - A.y[j], cz = A.z[k] - A.z[j];
    const double rbc = sqrt(mdg_dist(A.B, cx, cy, cz)), rbct = 1. / rbc;
    const double cost = (ax * cx + ay * cy + az * cz) / (rab * rbc);
    const double sint = fmax(1.e-8, sqrt(1. - (cost * cost)));
    const double theta = acos(cost);
    const double dth = theta - t.theta0;
    const double e = 0.5 * t.k * (dth * dth);
    const double da = t.k * dth / sint;
    const double fxa = da * (cx * rbct - ax * cost * rabt) * rabt;
    const double fya = da * (cy * rbct - ay * cost * rabt) * rabt;
    const double fza = da * (cz * rbct - az * cost * rabt) * rabt;
    const double fxc = da * (ax * rabt - cx * cost * rbct) * rbct;
    const double fyc = da * (ay * rabt - cy * cost * rbct) * rbct;
    const double fzc = da * (az * rabt - cz * cost * rbct) * rbct;
    add_force(A, i, fxa, fya, fza);
    add_force(A, j, -fxa - fxc, -fya - fyc, -fza - fzc);
    add_force(A, k, fxc, fyc, fzc);
    return e;
}

// internal.c:247-425 (dihedral) = 427-604 (improper)
__device__ double dihe_term(const BondedArgs &A, const MDG_DIHE &t)
{
#pragma clang fp contract(off)
    const int i = t.a, j = t.b, k = t.c, l = t.d;
    double ab[3] = {A.x[i] - A.x[j], A.y[i] - A.y[j], A.z[i] - A.z[j]};
    mdg_dist(A.B, ab[0], ab[1], ab[2]);
    double bc[3] = {A.x[j] - A.x[k], A.y[j] - A.y[k], A.z[j] - A.z[k]};
    const double rbc = sqrt(mdg_dist(A.B, bc[0], bc[1], bc[2]));
    double cd[3] = {A.x[k] - A.x[l], A.y[k] - A.y[l], A.z[k] - A.z[l]};
    mdg_dist(A.B, cd[0], cd[1], cd[2]);
    const double pb[3] = {ab[1] * bc[2] - ab[2] * bc[1], ab[2] * bc[0] - ab[0] * bc[2],
                          ab[0] * bc[1] - ab[1] * bc[0]};
    const double r2pb = pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2], rpb = sqrt(r2pb);
    const double pc[3] = {bc[1] * cd[2] - bc[2] * cd[1], bc[2] * cd[0] - bc[0] * cd[2],
                          bc[0] * cd[1] - bc[1] * cd[0]};
    const double r2pc = pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2], rpc = sqrt(r2pc);
    const double pbpc = pb[0] * pc[0] + pb[1] * pc[1] + pb[2] * pc[2];
    const double cosp = pbpc / (rpb * rpc);
    double sinp = (bc[0] * (pc[1] * pb[2] - pc[2] * pb[1]) + bc[1] * (pb[0] * pc[2] - pb[2] * pc[0]) +
                   bc[2] * (pc[0] * pb[1] - pc[1] * pb[0])) / (rpb * rpc * rbc);
    double phi = atan2(sinp, cosp);
    const double DBL_EPS = 2.220446049250313e-16;
    sinp = sinp >= 0. ? fmax(DBL_EPS, fabs(sinp)) : -fmax(DBL_EPS, fabs(sinp));
    double e, dd;
    if (t.type == 2) {
        const double twopi = 2. * 3.141592653589793;
        phi = phi - t.phi0;
        phi = phi - mdg_nint(phi / twopi) * twopi;
        e = 0.5 * t.k * (phi * phi);
        dd = t.k * phi / (rpb * rpc * sinp);
    } else {
        e = t.k * (1. + cos(t.mult * phi - t.phi0));
        dd = -t.k * t.mult * sin(t.mult * phi - t.phi0) / (rpb * rpc * sinp);
    }
    double fa[3], fb[3], fc[3], fd[3];
    fa[0] = dd * ((-pc[1] * bc[2] + pc[2] * bc[1]) - pbpc * (-pb[1] * bc[2] + pb[2] * bc[1]) / r2pb);
    fa[1] = dd * ((pc[0] * bc[2] - pc[2] * bc[0]) - pbpc * (pb[0] * bc[2] - pb[2] * bc[0]) / r2pb);
    fa[2] = dd * ((-pc[0] * bc[1] + pc[1] * bc[0]) - pbpc * (-pb[0] * bc[1] + pb[1] * bc[0]) / r2pb);
    fc[0] = dd * ((-pc[1] * ab[2] + pc[2] * ab[1]) - pbpc * (-pb[1] * ab[2] + pb[2] * ab[1]) / r2pb);
    fc[1] = dd * ((pc[0] * ab[2] - pc[2] * ab[0]) - pbpc * (pb[0] * ab[2] - pb[2] * ab[0]) / r2pb);
    fc[2] = dd * ((-pc[0] * ab[1] + pc[1] * ab[0]) - pbpc * (-pb[0] * ab[1] + pb[1] * ab[0]) / r2pb);
    fb[0] = dd * ((-pb[1] * cd[2] + pb[2] * cd[1]) - pbpc * (-pc[1] * cd[2] + pc[2] * cd[1]) / r2pc);
    fb[1] = dd * ((pb[0] * cd[2] - pb[2] * cd[0]) - pbpc * (pc[0] * cd[2] - pc[2] * cd[0]) / r2pc);
    fb[2] = dd * ((-pb[0] * cd[1] + pb[1] * cd[0]) - pbpc * (-pc[0] * cd[1] + pc[1] * cd[0]) / r2pc);
    fd[0] = dd * ((-pb[1] * bc[2] + pb[2] * bc[1]) - pbpc * (-pc[1] * bc[2] + pc[2] * bc[1]) / r2pc);
    fd[1] = dd * ((pb[0] * bc[2] - pb[2] * bc[0]) - pbpc * (pc[0] * bc[2] - pc[2] * bc[0]) / r2pc);
    fd[2] = dd * ((-pb[0] * bc[1] + pb[1] * bc[0]) - pbpc * (-pc[0] * bc[1] + pc[1] * bc[0]) / r2pc);
    add_force(A, i, fa[0], fa[1], fa[2]);
    add_force(A, j, -fa[0] - fc[0] + fb[0], -fa[1] - fc[1] + fb[1], -fa[2] - fc[2] + fb[2]);
    add_force(A, k, fc[0] - fb[0] - fd[0], fc[1] - fb[1] - fd[1], fc[2] - fb[2] - fd[2]);
    add_force(A, l, fd[0], fd[1], fd[2]);
    return e;
}

// coulomb14_shift1 / shift2 / switch (elec.c:334-465); NOELEC and FULL
// select coulomb14_none (energy.c:37-54).
__device__ double coulomb14(const MdgParams &P, int type, double scal14, double qel, double r2,
                            double rt, double &de)
{
#pragma clang fp contract(off)
    de = 0.;
    if (type == 2 || type == 3) {
        if (r2 > P.cutOff2) return 0.;
        double s1, s2, sf, dsf;
        if (type == 2) {
            s1 = rt * r2 * P.rcutOff;
            s2 = r2 * P.rcutOff2;
            sf = 1. - 2. * s1 + s2;
            dsf = 2. * (s2 - s1);
        } else {
            s1 = r2 * P.rcutOff2;
            s2 = s1 * s1;
            sf = 1. - 2. * s1 + s2;
            dsf = 4. * (s2 - s1);
        }
        const double pe = scal14 * P.chargeConst * qel * rt;
        de = pe * rt * (dsf - sf);
        return pe * sf;
    }
    if (type == 4) {
        const double pe = scal14 * P.chargeConst * qel * rt;
        if (r2 <= P.cutOn2) {
            de = -pe * rt;
            return pe;
        }
        if (r2 > P.cutOff2) return 0.;
        const double sw = P.cutOff2 - r2;
        const double sf = (sw * sw) * (P.cutOff2 + 2. * r2 - 3. * P.cutOn2) * P.switch2;
        const double dsf = 12. * r2 * sw * (P.cutOn2 - r2) * P.switch2;
        de = pe * rt * (dsf - sf);
        return pe * sf;
    }
    return 0.;
}

// vdw14_switch (vdw.c:270-311)
__device__ double vdw14(const MdgParams &P, double scal14, double veps, double vsig, double r2,
                        double rt, double &dv)
{
#pragma clang fp contract(off)
    dv = 0.;
    double s6 = vsig * rt;
    s6 = s6 * s6 * s6 * s6 * s6 * s6;
    const double s12 = s6 * s6;
    if (r2 <= P.cutOn2) {
        dv = 24. * scal14 * veps * rt * (s6 - 2. * s12);
        return 4. * scal14 * veps * (s12 - s6);
    }
    if (r2 > P.cutOff2) return 0.;
    const double sw = P.cutOff2 - r2;
    const double sf = (sw * sw) * (P.cutOff2 + 2. * r2 - 3. * P.cutOn2) * P.switch2;
    const double dsf = 12. * r2 * sw * (P.cutOn2 - r2) * P.switch2;
    const double pv = 4. * scal14 * veps * (s12 - s6);
    const double dpv = 24. * scal14 * veps * (s6 - 2. * s12);
    dv = (pv * dsf + dpv * sf) * rt;
    return pv * sf;
}

// One 1-4 pair (energy.c:371-419 / 670-724); e[0..3] += elec, vdw,
// virelec, virvdw.
__device__ void pair14(const BondedArgs &A, int k, double (&e)[4])
{
#pragma clang fp contract(off)
    const int i = A.l14[2 * k], j = A.l14[2 * k + 1];
    double dx = A.x[j] - A.x[i], dy = A.y[j] - A.y[i], dz = A.z[j] - A.z[i];
    const double r2 = mdg_dist(A.B, dx, dy, dz);
    const double r = sqrt(r2), rt = 1. / r;
    double dv = 0.;
    const double ev = A.vdw14 ? vdw14(A.P, A.scal14, A.eps14[i] * A.eps14[j],
                                      A.sig14[i] + A.sig14[j], r2, rt, dv)
                              : 0.;
    double fs;
    if (A.ewald14) {
        const double qel = A.P.chargeConst * A.q[i] * A.q[j];
        const double ar = A.P.alpha * r;
        const double ed = A.scal14 * qel * erfc(ar) * rt;
        double dd = 2.0 * A.scal14 * qel * A.P.alpha * exp(-(ar * ar)) / SQRTPI;
        dd = -(ed + dd) * rt;
        const double ec = (A.scal14 - 1.0) * qel * erf(ar) * rt;
        double dc = 2.0 * (A.scal14 - 1.0) * qel * A.P.alpha * exp(-(ar * ar)) / SQRTPI;
        dc = (dc - ec) * rt;
        e[0] += ed + ec;
        e[2] += dd * r + dc * r;
        fs = dd + dc + dv;
    } else {
        double de;
        e[0] += coulomb14(A.P, A.elec14, A.scal14, A.q[i] * A.q[j], r2, rt, de);
        e[2] += de * r;
        fs = de + dv;
    }
    e[1] += ev;
    e[3] += dv * r;
    const double fxj = fs * dx * rt, fyj = fs * dy * rt, fzj = fs * dz * rt;
    add_force(A, i, fxj, fyj, fzj);
    add_force(A, j, -fxj, -fyj, -fzj);
}

#define BD_NV 11   // bond, ub, ang, dihe, impr, virbond, virub, elec14, vdw14, virelec14, virvdw14

__global__ __launch_bounds__(256) void k_bonded(BondedArgs A)
{
    double acc[BD_NV];
#pragma unroll
    for (int k = 0; k < BD_NV; k++) acc[k] = 0.;
    const long o1 = A.nBond, o2 = o1 + A.nAngle, o3 = o2 + A.nUb, o4 = o3 + A.nDihe,
               o5 = o4 + A.nImpr, o6 = o5 + A.n14;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < o6;
         t += (long)gridDim.x * blockDim.x) {
        if (t < o1) {
            acc[0] += bond_term(A, A.bond[t], true, acc[5]);
        } else if (t < o2) {
            acc[2] += angle_term(A, A.angle[t - o1]);
        } else if (t < o3) {
            acc[1] += bond_term(A, A.ub[t - o2], false, acc[6]);
        } else if (t < o4) {
            acc[3] += dihe_term(A, A.dihe[t - o3]);
        } else if (t < o5) {
            acc[4] += dihe_term(A, A.impr[t - o4]);
        } else {
            double e[4] = {0., 0., 0., 0.};
            pair14(A, (int)(t - o5), e);
            acc[7] += e[0];
            acc[8] += e[1];
            acc[9] += e[2];
            acc[10] += e[3];
        }
    }
    mdg_block_partials<BD_NV>(acc, A.part);
}

__global__ void k_fix_to_f64(int n, long long *__restrict__ fix, double *fx, double *fy, double *fz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = 1.0 / MDG_FSCALE;
    fx[i] += (double)fix[i] * s;
    fy[i] += (double)fix[n + i] * s;
    fz[i] += (double)fix[2 * (long)n + i] * s;
    fix[i] = 0;
    fix[n + i] = 0;
    fix[2 * (long)n + i] = 0;
}

template <class T>
int upload_terms(mdg_ctx *ctx, T **dev, int *cap, const T *host, int n)
{
    if (n <= 0) return MDG_OK;
    if (n > *cap) {
        if (*dev) HIPCHK(hipFree(*dev));
        HIPCHK(hipMalloc(dev, (size_t)n * sizeof(T)));
        *cap = n;
    }
    HIPCHK(hipMemcpyAsync(*dev, host, (size_t)n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return MDG_OK;
}

}  // namespace

// Term tables of the context's bonded topology (device copies).
int mdg_bonded_upload(mdg_ctx *ctx, int nBond, const MDG_BOND *bond, int nUb, const MDG_BOND *ub,
                      int nAngle, const MDG_ANGLE *angle, int nDihe, const MDG_DIHE *dihe,
                      int nImpr, const MDG_DIHE *impr, int n14, const int *list14,
                      const double *eps14, const double *sig14)
{
    // every atom index a kernel will dereference lies in [0, n)
    const int n = ctx->n;
    auto bad = [n](int a) { return a < 0 || a >= n; };
    for (int i = 0; i < nBond; i++)
        if (bad(bond[i].a) || bad(bond[i].b)) return MDG_EARG;
    for (int i = 0; i < nUb; i++)
        if (bad(ub[i].a) || bad(ub[i].b)) return MDG_EARG;
    for (int i = 0; i < nAngle; i++)
        if (bad(angle[i].a) || bad(angle[i].b) || bad(angle[i].c)) return MDG_EARG;
    for (int i = 0; i < nDihe; i++)
        if (bad(dihe[i].a) || bad(dihe[i].b) || bad(dihe[i].c) || bad(dihe[i].d)) return MDG_EARG;
    for (int i = 0; i < nImpr; i++)
        if (bad(impr[i].a) || bad(impr[i].b) || bad(impr[i].c) || bad(impr[i].d)) return MDG_EARG;
    for (int i = 0; i < 2 * n14; i++)
        if (bad(list14[i])) return MDG_EARG;
    MdgBonded &D = ctx->bd;
    RETCHK(upload_terms(ctx, &D.bond, &D.capBond, bond, nBond));
    RETCHK(upload_terms(ctx, &D.ub, &D.capUb, ub, nUb));
    RETCHK(upload_terms(ctx, &D.angle, &D.capAngle, angle, nAngle));
    RETCHK(upload_terms(ctx, &D.dihe, &D.capDihe, dihe, nDihe));
    RETCHK(upload_terms(ctx, &D.impr, &D.capImpr, impr, nImpr));
    RETCHK(upload_terms(ctx, &D.l14, &D.cap14, list14, 2 * n14));
    D.nBond = nBond;
    D.nUb = nUb;
    D.nAngle = nAngle;
    D.nDihe = nDihe;
    D.nImpr = nImpr;
    D.n14 = n14;
    if (n14 > 0 && eps14 && sig14) {
        if (!D.eps14) {
            HIPCHK(hipMalloc(&D.eps14, (size_t)ctx->n * sizeof(double)));
            HIPCHK(hipMalloc(&D.sig14, (size_t)ctx->n * sizeof(double)));
        }
        HIPCHK(hipMemcpyAsync(D.eps14, eps14, (size_t)ctx->n * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(D.sig14, sig14, (size_t)ctx->n * sizeof(double),
                              hipMemcpyHostToDevice, ctx->stream));
    }
    return MDG_OK;
}

// Bonded terms and 1-4 pairs of the current positions: forces into
// fx/fy/fz (FP64 atomics) or, with fix, into the int64 per-atom
// accumulators; block partials into ctx->epart + PART_BOND and their
// BD_NV sums into ctx->scal + S_BONDED.
int mdg_bonded_blocks(const mdg_ctx *ctx)
{
    const MdgBonded &D = ctx->bd;
    const long total = (long)D.nBond + D.nAngle + D.nUb + D.nDihe + D.nImpr + D.n14;
    return std::max(1, std::min(nblk(total, 256), 1024));
}

int mdg_bonded_forces(mdg_ctx *ctx, double *fx, double *fy, double *fz, long long *fix,
                      bool reduce)
{
    const MdgBonded &D = ctx->bd;
    const long total = (long)D.nBond + D.nAngle + D.nUb + D.nDihe + D.nImpr + D.n14;
    if (total == 0 && reduce) return mdg_fill_f64(ctx, ctx->scal + S_BONDED, BD_NV, 0.);
    BondedArgs A{};
    A.P = ctx->P;
    A.B = ctx->B;
    A.x = ctx->x;
    A.y = ctx->y;
    A.z = ctx->z;
    A.q = ctx->q;
    A.eps14 = D.eps14;
    A.sig14 = D.sig14;
    A.bond = D.bond;
    A.ub = D.ub;
    A.angle = D.angle;
    A.dihe = D.dihe;
    A.impr = D.impr;
    A.l14 = D.l14;
    A.nBond = D.nBond;
    A.nAngle = D.nAngle;
    A.nUb = D.nUb;
    A.nDihe = D.nDihe;
    A.nImpr = D.nImpr;
    A.n14 = D.n14;
    A.scal14 = D.scal14;
    A.ewald14 = D.ewald14;
    A.elec14 = D.elec14;
    A.vdw14 = D.vdw14 && D.eps14;
    A.fx = fx;
    A.fy = fy;
    A.fz = fz;
    A.fix = fix;
    A.n = ctx->n;
    A.part = ctx->epart + PART_BOND;
    const int nb = mdg_bonded_blocks(ctx);
    hipLaunchKernelGGL(k_bonded, dim3(nb), dim3(256), 0, ctx->stream, A);
    LAUNCHCHK();
    if (!reduce) return MDG_OK;
    return mdg_reduce_partials(ctx, ctx->epart + PART_BOND, nb, BD_NV, ctx->scal + S_BONDED);
}

// The context's bonded topology for the device-resident step: term tables,
// the 1-4 settings, and the exclusion table rebuilt from every term
// (list.c:108-894 / 1125-1791).  No reference counterpart as one call (the
// reference passes the tables to energy() every step).
extern "C" int mdg_set_bonded(mdg_ctx *ctx, const MDG_CTRL *ctrl, const MDG_PARAM *param,
                              const MDG_BOND bond[], const MDG_BOND ub[], const MDG_ANGLE angle[],
                              const MDG_DIHE dihe[], const MDG_DIHE impr[], const double eps14[],
                              const double sig14[])
{
    if (!ctx || !ctrl || !param || param->nAtom != ctx->n) return MDG_EARG;
    if (ctx->bt) {
        ctx->err = "bonded terms are not available in a replica batch";
        return MDG_EUNSUPPORTED;
    }
    if ((param->nBond > 0 && !bond) || (param->nUb > 0 && !ub) || (param->nAngle > 0 && !angle) ||
        (param->nDihedral > 0 && !dihe) || (param->nImproper > 0 && !impr))
        return MDG_EARG;
    if (ctrl->keyNb14 && (!eps14 || !sig14)) return MDG_EARG;
    std::vector<int> l14;
    RETCHK(mdg_build_exclusions(ctx, param->nBond, bond,
                                ctrl->keyConstH ? (int)ctx->hCons.size() : 0, ctx->hCons.data(),
                                param->nAngle, angle, param->nDihedral, dihe, param->nImproper,
                                impr, &l14));
    MdgBonded &D = ctx->bd;
    D.scal14 = param->scal14;
    D.ewald14 = ctx->P.ewald != 0;   // ewald14_energy, else nonbond14_energy (energy.c:151-165)
    D.elec14 = ctrl->elecType;
    D.vdw14 = ctrl->vdwType == 2;
    const int n14 = ctrl->keyNb14 ? (int)(l14.size() / 2) : 0;
    RETCHK(mdg_bonded_upload(ctx, param->nBond, bond, param->nUb, ub, param->nAngle, angle,
                             param->nDihedral, dihe, param->nImproper, impr, n14, l14.data(),
                             eps14, sig14));
    if (!D.fix) {
        HIPCHK(hipMalloc(&D.fix, 3 * (size_t)ctx->n * sizeof(long long)));
        RETCHK(mdg_fill_i64(ctx, D.fix, 3 * (long)ctx->n));
    }
    D.on = (long)D.nBond + D.nUb + D.nAngle + D.nDihe + D.nImpr + D.n14 > 0;
    // captured step graphs hold the old record and force-assembly arguments
    mdg_invalidate_graphs(ctx);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDG_OK;
}

// Adds the int64 per-atom accumulators fix into fx/fy/fz and clears them.
int mdg_fix_to_f64(mdg_ctx *ctx, long long *fix, double *fx, double *fy, double *fz)
{
    hipLaunchKernelGGL(k_fix_to_f64, dim3(nblk(ctx->n, 256)), dim3(256), 0, ctx->stream, ctx->n,
                       fix, fx, fy, fz);
    LAUNCHCHK();
    return MDG_OK;
}

void mdg_bonded_free(mdg_ctx *ctx)
{
    MdgBonded &D = ctx->bd;
    void *p[] = {D.bond, D.ub, D.angle, D.dihe, D.impr, D.l14, D.eps14, D.sig14, D.fix};
    for (void *q : p)
        if (q) (void)hipFree(q);
    D = MdgBonded{};
}
