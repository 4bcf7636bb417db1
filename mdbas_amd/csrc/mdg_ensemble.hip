// mdg_ensemble.hip -- the thermostat and barostat integrators behind the
// drop-in vv_integrate: Berendsen NVT and NPT, Nose-Hoover NVT and NPT
// (vv_nvt_b, vv_npt_b, vv_nvt_h, vv_npt_h, integrate.c:1636-2618).
//
// Every per-atom operation runs on the device, on the context's state:
// kicks, drifts and SHAKE through the generic-group VV kernels (with the
// VvOpts switches), velocity scaling, and the mass-weighted sums (kinetic
// energy and stress, centre of mass, mean velocity) as block partials
// reduced in a fixed order.  What stays on the host is the bath itself: a
// few doubles (chiT, chiP, the volume, conint, consv, the box vectors)
// advanced between launches in the reference's expression order.
//
// Two reference behaviours are kept as they are:
//  * the integrators' virshake is a local that vv_npt_b (without
//    constraints) and vv_npt_h (its first Nose cycle, and all of stage 2)
//    read before a SHAKE call has set it; it reads zero here;
//  * vv_npt_h's second Nose cycle restarts the bath from the step's start
//    but continues from the first cycle's positions and velocities
//    (integrate.c:2415-2437 saves them where vv_npt_b restores them).
#include <cmath>

#include "mdg_shake.h"

namespace {

constexpr double WATERCOMP = 0.007372;   // global.h:56
constexpr double RBOLTZUI = 0.83144621;  // global.h:47
constexpr int HOOVER = 5;                // integrate.c:2176

// The constraint vectors at the call's entry (integrate.c:1790-1810),
// minimum image in the entry box.
__global__ void k_cons_dd(int nc, const int *__restrict__ ca, const int *__restrict__ cb,
                          const double *__restrict__ x, const double *__restrict__ y,
                          const double *__restrict__ z, MdgBox B, double *__restrict__ dd)
{
#pragma clang fp contract(off)
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double dx = x[cb[c]] - x[ca[c]], dy = y[cb[c]] - y[ca[c]], dz = z[cb[c]] - z[ca[c]];
    mdg_dist(B, dx, dy, dz);
    dd[c] = dx;
    dd[nc + c] = dy;
    dd[2 * nc + c] = dz;
}

// Per-block partials of m v^2 and the kinetic stress (kinetic,
// utils.c:502-515; stress_kinetic, utils.c:517-542).
__global__ void __launch_bounds__(256)
k_kin_parts(int n, const double *__restrict__ vx, const double *__restrict__ vy,
            const double *__restrict__ vz, const double *__restrict__ m, double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const double u = vx[i], v = vy[i], w = vz[i], mi = m[i];
        a[0] = mi * ((u * u + v * v) + w * w);
        a[1] = mi * (u * u);
        a[2] = mi * u * v;
        a[3] = mi * u * w;
        a[4] = mi * (v * v);
        a[5] = mi * v * w;
        a[6] = mi * (w * w);
    }
    mdg_block_partials<7>(a, part);
}

// Per-block partials of sum m a, m b, m c (centre of mass, mean momentum).
__global__ void __launch_bounds__(256)
k_mass_parts(int n, const double *__restrict__ m, const double *__restrict__ a,
             const double *__restrict__ b, const double *__restrict__ c, double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s[3] = {0, 0, 0};
    if (i < n) {
        s[0] = m[i] * a[i];
        s[1] = m[i] * b[i];
        s[2] = m[i] * c[i];
    }
    mdg_block_partials<3>(s, part);
}

__global__ void k_vscale(int n, double *__restrict__ vx, double *__restrict__ vy,
                         double *__restrict__ vz, double f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] *= f;
    vy[i] *= f;
    vz[i] *= f;
}

__global__ void k_vshift(int n, double *__restrict__ vx, double *__restrict__ vy,
                         double *__restrict__ vz, double c0, double c1, double c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] -= c0;
    vy[i] -= c1;
    vz[i] -= c2;
}

// Half kick, integrate.c:1554-1561.
__global__ void k_kick(int n, double dt, double *__restrict__ vx, double *__restrict__ vy,
                       double *__restrict__ vz, const double *__restrict__ fx,
                       const double *__restrict__ fy, const double *__restrict__ fz,
                       const double *__restrict__ rmass)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] += 0.5 * dt * fx[i] * rmass[i];
    vy[i] += 0.5 * dt * fy[i] * rmass[i];
    vz[i] += 0.5 * dt * fz[i] * rmass[i];
}

// image_update (utils.c:162-218) in the given box.
__global__ void k_wrap(int n, double *__restrict__ x, double *__restrict__ y,
                       double *__restrict__ z, MdgBox B)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xs = x[i], ys = y[i], zs = z[i];
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
    x[i] = xs;
    y[i] = ys;
    z[i] = zs;
}

// The bath of one call: the reference's BATH, PARAM targets and ENERGY
// members, with the integrator's virshake local.
struct Bath {
    double tauT, tauP, chiT, chiP, kinTemp0, temp0, press0;
    double virpot, conint, consv, kin;
    double vs = 0.;
};

struct Ens {
    mdg_ctx *ctx;
    int n;
    double dt;

    int blocks() const { return nblk(n, 256); }

    int kinetic(double *kin, double st[6] = nullptr)
    {
        double *part = ctx->epart + PART_MISC;
        hipLaunchKernelGGL(k_kin_parts, dim3(blocks()), dim3(256), 0, ctx->stream, n, ctx->vx,
                           ctx->vy, ctx->vz, ctx->mass, part);
        LAUNCHCHK();
        double s[7];
        RETCHK(mdg_reduce_partials(ctx, part, blocks(), 7, ctx->scal + S_ENS));
        HIPCHK(hipMemcpyAsync(s, ctx->scal + S_ENS, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *kin = s[0] * 0.5;
        if (st)
            for (int k = 0; k < 6; k++) st[k] = s[1 + k];
        return MDG_OK;
    }

    int mass_sum(const double *a, const double *b, const double *c, double out[3])
    {
        double *part = ctx->epart + PART_MISC;
        hipLaunchKernelGGL(k_mass_parts, dim3(blocks()), dim3(256), 0, ctx->stream, n, ctx->mass,
                           a, b, c, part);
        LAUNCHCHK();
        RETCHK(mdg_reduce_partials(ctx, part, blocks(), 3, ctx->scal + S_ENS));
        HIPCHK(hipMemcpyAsync(out, ctx->scal + S_ENS, 3 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return MDG_OK;
    }

    int vscale(double f)
    {
        hipLaunchKernelGGL(k_vscale, dim3(blocks()), dim3(256), 0, ctx->stream, n, ctx->vx,
                           ctx->vy, ctx->vz, f);
        LAUNCHCHK();
        return MDG_OK;
    }

    int kick()
    {
        hipLaunchKernelGGL(k_kick, dim3(blocks()), dim3(256), 0, ctx->stream, n, dt, ctx->vx,
                           ctx->vy, ctx->vz, ctx->fx, ctx->fy, ctx->fz, ctx->rmass);
        LAUNCHCHK();
        return MDG_OK;
    }

    int cons_dd()
    {
        if (ctx->nConst <= 0) return MDG_OK;
        hipLaunchKernelGGL(k_cons_dd, dim3(nblk(ctx->nConst, 256)), dim3(256), 0, ctx->stream,
                           ctx->nConst, ctx->ca, ctx->cb, ctx->x, ctx->y, ctx->z, ctx->B, ctx->dd);
        LAUNCHCHK();
        return MDG_OK;
    }

    // positions and velocities into / out of ensSave
    int save(bool store)
    {
        double *state[6] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz};
        for (int k = 0; k < 6; k++) {
            double *s = ctx->ensSave + (size_t)k * n;
            HIPCHK(hipMemcpyAsync(store ? s : state[k], store ? state[k] : s, n * sizeof(double),
                                  hipMemcpyDeviceToDevice, ctx->stream));
        }
        return MDG_OK;
    }

    int stage(int st, const VvOpts &o, double *kin7, double *vs)
    {
        return mdg_vv_stage(ctx, st, ctx->fx, ctx->fy, ctx->fz, kin7, vs, true, nullptr, &o);
    }

    // one thermostat-barostat sub-step of vv_npt_h (integrate.c:2296-2382)
    int hoover(Bath &b, double qmass, double pmass, double cqts, double chts, double &volume)
    {
        for (int half = 0; half < 2; half++) {
            RETCHK(kinetic(&b.kin));
            b.chiT += 0.5 * cqts *
                      (2.0 * (b.kin - b.kinTemp0) + pmass * (b.chiP * b.chiP) - RBOLTZUI * b.temp0) /
                      qmass;
            const double l = exp(-cqts * b.chiT);
            RETCHK(vscale(l));
            b.kin *= l * l;
            b.conint += cqts * b.chiT * (RBOLTZUI * b.temp0 + qmass / (b.tauT * b.tauT));
            b.chiT += 0.5 * cqts *
                      (2.0 * (b.kin - b.kinTemp0) + pmass * (b.chiP * b.chiP) - RBOLTZUI * b.temp0) /
                      qmass;
            if (half == 1) break;
            // barostat between the two thermostat halves
            b.chiP += 0.5 * chts *
                      (((2.0 * b.kin - b.virpot - b.vs) - 3.0 * b.press0 * volume) / pmass -
                       b.chiP * b.chiT);
            const double g = exp(-chts * b.chiP);
            RETCHK(vscale(g));
            b.kin *= g * g;
            volume *= exp(3.0 * chts * b.chiP);
            b.chiP += 0.5 * chts *
                      (((2.0 * b.kin - b.virpot - b.vs) - 3.0 * b.press0 * volume) / pmass -
                       b.chiP * b.chiT);
        }
        return MDG_OK;
    }
};

// scale_box / vv_scale_box (utils.c:291-459): new vectors, everything
// else derived as init_box does.
void rescale_box(MDG_PBC *box, const double cell0[9], double s)
{
#pragma clang fp contract(off)
    box->a1 = s * cell0[0]; box->a2 = s * cell0[1]; box->a3 = s * cell0[2];
    box->b1 = s * cell0[3]; box->b2 = s * cell0[4]; box->b3 = s * cell0[5];
    box->c1 = s * cell0[6]; box->c2 = s * cell0[7]; box->c3 = s * cell0[8];
    box->a = sqrt(box->a1 * box->a1 + box->a2 * box->a2 + box->a3 * box->a3);
    box->b = sqrt(box->b1 * box->b1 + box->b2 * box->b2 + box->b3 * box->b3);
    box->c = sqrt(box->c1 * box->c1 + box->c2 * box->c2 + box->c3 * box->c3);
    MdgBox B;
    mdg_box_from(box, B);
    box->u1 = B.u1; box->u2 = B.u2; box->u3 = B.u3;
    box->v1 = B.v1; box->v2 = B.v2; box->v3 = B.v3;
    box->w1 = B.w1; box->w2 = B.w2; box->w3 = B.w3;
    box->pa = B.pa; box->pb = B.pb; box->pc = B.pc;
    box->u = B.u; box->v = B.v; box->w = B.w;
    box->det = B.det;
    box->vol = B.vol;
}

void cell_of(const MDG_PBC *box, double c[9])
{
    const double v[9] = {box->a1, box->a2, box->a3, box->b1, box->b2,
                         box->b3, box->c1, box->c2, box->c3};
    for (int k = 0; k < 9; k++) c[k] = v[k];
}

void add_stress(MDG_PBC *box, const double s[6])
{
    // integrate.c:1603-1611 (the SHAKE stress of stage 2 is zero)
    box->stress1 += s[0];
    box->stress2 += s[1];
    box->stress3 += s[2];
    box->stress4 += s[1];
    box->stress5 += s[3];
    box->stress6 += s[4];
    box->stress7 += s[2];
    box->stress8 += s[4];
    box->stress9 += s[5];
}

}  // namespace

// One stage of ensemble ens (1 NVT_B, 2 NPT_B, 3 NVT_H, 4 NPT_H) on the
// context's device state (positions, velocities, forces, masses already
// uploaded); ener, bath and (NPT) box are updated as the reference does.
int mdg_vv_ensemble(mdg_ctx *ctx, const MDG_PARAM *param, MDG_ENERGY *ener, MDG_PBC *box,
                    MDG_BATH *bath, int ens, int stage)
{
#pragma clang fp contract(off)
    if (!bath || !box || ens < 1 || ens > 4) return MDG_EARG;
    if (ctx->bt) {
        ctx->err = "thermostats run on single systems";
        return MDG_EUNSUPPORTED;
    }
    Ens E{ctx, ctx->n, param->timeStep};
    Bath b;
    b.tauT = bath->tauT;
    b.tauP = bath->tauP;
    b.chiT = bath->chiT;
    b.chiP = bath->chiP;
    b.kinTemp0 = param->kinTemp0;
    b.temp0 = param->temp0;
    b.press0 = param->press0;
    b.virpot = ener->virpot;
    b.conint = ener->conint;
    b.consv = ener->consv;
    b.kin = ener->kin;
    const bool shake = ctx->nConst > 0;
    const double dt = param->timeStep;
    VvOpts plain;
    double k7[8] = {0}, st[6];

    if (ens == 1) {                                       // vv_nvt_b, integrate.c:1636-1768
        if (stage == 1) {
            RETCHK(E.stage(1, plain, nullptr, &b.vs));
            if (shake) ener->virshake = b.vs;
        } else {
            RETCHK(E.stage(2, plain, k7, nullptr));
            b.kin = k7[0];
            const double l = sqrt(1.0 + dt / b.tauT * (b.kinTemp0 / b.kin - 1.0));
            RETCHK(E.vscale(l));
            b.kin *= l * l;
            double k;
            RETCHK(E.kinetic(&k, st));
            add_stress(box, st);
        }
    } else if (ens == 3) {                                // vv_nvt_h, integrate.c:1998-2169
        const double qmass = 2.0 * b.kinTemp0 * (b.tauT * b.tauT);
        if (stage == 1) {
            RETCHK(E.kinetic(&b.kin));
            b.chiT += 0.5 * dt * (b.kin - b.kinTemp0) / qmass;
            const double l = exp(-0.5 * dt * b.chiT);
            RETCHK(E.vscale(l));
            b.kin *= l * l;
            b.conint += 0.5 * dt * b.chiT * qmass / (b.tauT * b.tauT);
            b.chiT += 0.5 * dt * (b.kin - b.kinTemp0) / qmass;
            RETCHK(E.stage(1, plain, nullptr, &b.vs));
            if (shake) ener->virshake = b.vs;
        } else {
            RETCHK(E.stage(2, plain, k7, nullptr));
            b.kin = k7[0];
            b.chiT += 0.5 * dt * (b.kin - b.kinTemp0) / qmass;
            const double l = exp(-0.5 * dt * b.chiT);
            RETCHK(E.vscale(l));
            b.kin *= l * l;
            b.conint += 0.5 * dt * b.chiT * qmass / (b.tauT * b.tauT);
            b.chiT += 0.5 * dt * (b.kin - b.kinTemp0) / qmass;
            b.consv = b.conint + 0.5 * qmass * (b.chiT * b.chiT);
            double k;
            RETCHK(E.kinetic(&k, st));
            add_stress(box, st);
        }
    } else if (ens == 2) {                                // vv_npt_b, integrate.c:1770-1996
        const double volume = box->vol;
        if (stage == 1) {
            RETCHK(E.kinetic(&b.kin));
            RETCHK(E.cons_dd());
            RETCHK(E.kick());
            const int nose = shake ? 2 : 1;
            if (nose > 1) {
                if (!ctx->ensSave) HIPCHK(hipMalloc(&ctx->ensSave, (size_t)6 * E.n * sizeof(double)));
                RETCHK(E.save(true));
            }
            for (int k = 0; k < nose; k++) {
                VvOpts o;
                o.ddStored = 1;
                o.noKick = 1;
                o.npt = 1;
                if (k == nose - 1) {
                    const double pp = (2. * b.kin - b.virpot - b.vs) / (3. * volume);
                    const double g = 1. + WATERCOMP * dt * (pp - b.press0) / b.tauP;
                    o.dscale = cbrt(g);
                    double cell[9];
                    cell_of(box, cell);
                    rescale_box(box, cell, o.dscale);
                    mdg_box_from(box, ctx->B);   // SHAKE's images in the new box
                }
                RETCHK(E.stage(1, o, nullptr, &b.vs));
                if (shake) ener->virshake = b.vs;
                if (k < nose - 1) RETCHK(E.save(false));
            }
        } else {
            RETCHK(E.kick());
            RETCHK(E.kinetic(&b.kin));
            const double l = sqrt(1.0 + dt / b.tauT * (b.kinTemp0 / b.kin - 1.0));
            RETCHK(E.vscale(l));
            VvOpts o;
            o.noKick = 1;
            RETCHK(E.stage(2, o, k7, nullptr));
            b.kin = k7[0];
            add_stress(box, k7 + 1);
        }
    } else {                                              // vv_npt_h, integrate.c:2171-2618
        const double hts = 0.5 * dt;
        const double chts = hts / (double)HOOVER;
        const double cqts = 0.25 * dt / (double)HOOVER;
        double cell0[9];
        cell_of(box, cell0);
        double volume = box->vol;
        const double volume0 = box->vol;
        // total mass, summed in atom order as integrate.c:2214-2216 does
        std::vector<double> m(E.n);
        HIPCHK(hipMemcpyAsync(m.data(), ctx->mass, E.n * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        double masst = 0.;
        for (int i = 0; i < E.n; i++) masst += m[i];
        const double qmass = 2.0 * b.kinTemp0 * (b.tauT * b.tauT);
        const double pmass = 2.0 * b.kinTemp0 * (b.tauP * b.tauP);
        if (stage == 1) {
            const double chiT0 = b.chiT, chiP0 = b.chiP, cons0 = b.conint;
            RETCHK(E.cons_dd());
            const int nose = shake ? 2 : 1;
            for (int k = 0; k < nose; k++) {
                for (int kk = 0; kk < HOOVER; kk++) RETCHK(E.hoover(b, qmass, pmass, cqts, chts, volume));
                rescale_box(box, cell0, cbrt(volume / volume0));
                mdg_box_from(box, ctx->B);
                double com[3];
                RETCHK(E.mass_sum(ctx->x, ctx->y, ctx->z, com));
                VvOpts o;
                o.ddStored = 1;
                o.npt = 1;
                o.dscale = exp(dt * b.chiP);
                for (int d = 0; d < 3; d++) o.com[d] = com[d] / masst;
                RETCHK(E.stage(1, o, nullptr, &b.vs));
                if (shake) ener->virshake = b.vs;
                if (k < nose - 1) {
                    volume = volume0;
                    b.chiT = chiT0;
                    b.chiP = chiP0;
                    b.conint = cons0;
                }
            }
        } else {
            VvOpts o;
            o.noWrap = 1;
            RETCHK(E.stage(2, o, k7, nullptr));
            for (int kk = 0; kk < HOOVER; kk++) RETCHK(E.hoover(b, qmass, pmass, cqts, chts, volume));
            double vom[3];
            RETCHK(E.mass_sum(ctx->vx, ctx->vy, ctx->vz, vom));
            for (int d = 0; d < 3; d++) vom[d] /= masst;
            hipLaunchKernelGGL(k_vshift, dim3(E.blocks()), dim3(256), 0, ctx->stream, E.n, ctx->vx,
                               ctx->vy, ctx->vz, vom[0], vom[1], vom[2]);
            LAUNCHCHK();
            rescale_box(box, cell0, cbrt(volume / volume0));
            mdg_box_from(box, ctx->B);
            b.consv = b.conint + b.press0 * volume +
                      0.5 * (qmass * (b.chiT * b.chiT) + pmass * (b.chiP * b.chiP));
            RETCHK(E.kinetic(&b.kin, st));
            add_stress(box, st);
            hipLaunchKernelGGL(k_wrap, dim3(E.blocks()), dim3(256), 0, ctx->stream, E.n, ctx->x,
                               ctx->y, ctx->z, ctx->B);
            LAUNCHCHK();
        }
    }
    bath->chiT = b.chiT;
    bath->chiP = b.chiP;
    ener->conint = b.conint;
    ener->consv = b.consv;
    ener->kin = b.kin;
    // a rescaled box: list counts, net-charge energy, influence function
    if (ens == 2 || ens == 4) RETCHK(mdg_refresh_box(ctx, box));
    return MDG_OK;
}

// ---------------------------------------------------------------- device-resident thermostat

namespace {

// bathDev layout
enum { B_CHIT = 0, B_CONINT = 1, B_LAMBDA = 2, B_KIN = 3, B_REC_CONINT = 4, B_REC_CONSV = 5,
       B_REC_CHIT = 6, B_WORDS = 8 };

struct ThermoArgs {
    int ens, stage, rows;
    const double *part;          // kinetic partials, [rows][7], sum m v^2 first
    double *bath;
    double dt, tauT, kinTemp0;
};

// One block: K = sum of the partials' m v^2 / 2 (fixed order), then the
// bath's scalar update of vv_nvt_b (integrate.c:1723-1737) or vv_nvt_h
// (integrate.c:2033-2062 at stage 1, 2116-2134 at stage 2), in the
// reference's expression order; lambda goes to bath[B_LAMBDA].
__global__ void __launch_bounds__(256) k_thermo(ThermoArgs T)
{
#pragma clang fp contract(off)
    __shared__ double red[4];
    double v = 0.;
    for (int r = threadIdx.x; r < T.rows; r += blockDim.x) v += T.part[(long)r * 7];
    v = mdg_wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double *b = T.bath;
    double kin = ((red[0] + red[1]) + red[2] + red[3]) * 0.5;
    double lam;
    if (T.ens == 1) {
        lam = sqrt(1.0 + T.dt / T.tauT * (T.kinTemp0 / kin - 1.0));
        kin *= lam * lam;
        // mdbas.c:246 clears consv every step; vv_nvt_b leaves it and conint
        b[B_REC_CONINT] = b[B_CONINT];
        b[B_REC_CONSV] = 0.;
        b[B_REC_CHIT] = b[B_CHIT];
    } else {
        const double qmass = 2.0 * T.kinTemp0 * (T.tauT * T.tauT);
        double chiT = b[B_CHIT], conint = b[B_CONINT];
        chiT += 0.5 * T.dt * (kin - T.kinTemp0) / qmass;
        lam = exp(-0.5 * T.dt * chiT);
        kin *= lam * lam;
        conint += 0.5 * T.dt * chiT * qmass / (T.tauT * T.tauT);
        chiT += 0.5 * T.dt * (kin - T.kinTemp0) / qmass;
        b[B_CHIT] = chiT;
        b[B_CONINT] = conint;
        if (T.stage == 2) {
            b[B_REC_CONINT] = conint;
            b[B_REC_CONSV] = conint + 0.5 * qmass * (chiT * chiT);
            b[B_REC_CHIT] = chiT;
        }
    }
    b[B_LAMBDA] = lam;
    b[B_KIN] = kin;
}

// v *= lambda (bath[B_LAMBDA]); with part, block partials of m v^2 and the
// kinetic stress of the scaled velocities in [gridDim.x][7] (the rows the
// step's record sums), grid-stride over the atoms.
__global__ void __launch_bounds__(256)
k_vscale_bath(int n, double *__restrict__ vx, double *__restrict__ vy, double *__restrict__ vz,
              const double *__restrict__ m, const double *__restrict__ bath, double *part)
{
#pragma clang fp contract(off)
    const double lam = bath[B_LAMBDA];
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double u = vx[i] * lam, v = vy[i] * lam, w = vz[i] * lam;
        vx[i] = u;
        vy[i] = v;
        vz[i] = w;
        if (part) {
            const double mi = m[i];
            a[0] += mi * ((u * u + v * v) + w * w);
            a[1] += mi * (u * u);
            a[2] += mi * u * v;
            a[3] += mi * u * w;
            a[4] += mi * (v * v);
            a[5] += mi * v * w;
            a[6] += mi * (w * w);
        }
    }
    if (part) mdg_block_partials<7>(a, part);
}

}  // namespace

int mdg_thermo_stage(mdg_ctx *ctx, int stage)
{
    if (ctx->ens != 1 && ctx->ens != 3) return MDG_OK;
    if (stage == 1 && ctx->ens != 3) return MDG_OK;   // Berendsen acts at stage 2 only
    const int n = ctx->n;
    ThermoArgs T;
    T.ens = ctx->ens;
    T.stage = stage;
    T.bath = ctx->bathDev;
    T.dt = ctx->P.timeStep;
    T.tauT = ctx->tauT;
    T.kinTemp0 = ctx->kinTemp0;
    if (stage == 1) {
        // the kinetic energy of the step's incoming velocities
        T.rows = nblk(n, 256);
        T.part = ctx->epart + PART_ENS;
        hipLaunchKernelGGL(k_kin_parts, dim3(T.rows), dim3(256), 0, ctx->stream, n, ctx->vx,
                           ctx->vy, ctx->vz, ctx->mass, ctx->epart + PART_ENS);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_thermo, dim3(1), dim3(256), 0, ctx->stream, T);
        hipLaunchKernelGGL(k_vscale_bath, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n,
                           ctx->vx, ctx->vy, ctx->vz, ctx->mass, ctx->bathDev, nullptr);
        LAUNCHCHK();
        return MDG_OK;
    }
    // after SHAKE-V: the stage's kinetic partials, then the scaled
    // velocities' partials in their place for the record
    T.rows = mdg_vv_blocks(ctx);
    T.part = ctx->epart + PART_KIN;
    hipLaunchKernelGGL(k_thermo, dim3(1), dim3(256), 0, ctx->stream, T);
    hipLaunchKernelGGL(k_vscale_bath, dim3(T.rows), dim3(256), 0, ctx->stream, n, ctx->vx, ctx->vy,
                       ctx->vz, ctx->mass, ctx->bathDev, ctx->epart + PART_KIN);
    LAUNCHCHK();
    return MDG_OK;
}

extern "C" int mdg_set_bath(mdg_ctx *ctx, const MDG_CTRL *ctrl, const MDG_PARAM *param,
                            const MDG_BATH *bath, const MDG_ENERGY *ener)
{
    if (!ctx || !ctrl || !param) return MDG_EARG;
    const int ens = (int)ctrl->ens;
    if (ens == 2 || ens == 4) {
        ctx->err = "the barostats change the box, which the device step holds fixed: run them "
                   "through mdg_vv_integrate";
        return MDG_EUNSUPPORTED;
    }
    if (ens < 0 || ens > 4) return MDG_EARG;
    if (ens != 0 && (!bath || !(bath->tauT > 0.) || !(param->kinTemp0 > 0.))) return MDG_EARG;
    if (ens != 0 && ctx->bt) {
        ctx->err = "the device thermostat runs on single systems";
        return MDG_EUNSUPPORTED;
    }
    if (!ctx->bathDev) HIPCHK(hipMalloc(&ctx->bathDev, B_WORDS * sizeof(double)));
    double b[B_WORDS] = {0};
    b[B_CHIT] = bath ? bath->chiT : 0.;
    b[B_CONINT] = ener ? ener->conint : 0.;
    b[B_LAMBDA] = 1.;
    b[B_REC_CONINT] = b[B_CONINT];
    b[B_REC_CONSV] = ener ? ener->consv : 0.;
    b[B_REC_CHIT] = b[B_CHIT];
    HIPCHK(hipMemcpyAsync(ctx->bathDev, b, sizeof b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->ens = ctx->ensCreate = ens;
    ctx->tauT = bath ? bath->tauT : 0.;
    ctx->kinTemp0 = param->kinTemp0;
    mdg_invalidate_graphs(ctx);
    return MDG_OK;
}

extern "C" int mdg_get_bath(mdg_ctx *ctx, MDG_BATH *bath, MDG_ENERGY *ener)
{
    if (!ctx) return MDG_EARG;
    double b[B_WORDS] = {0};
    if (ctx->ens && ctx->bathDev) {
        HIPCHK(hipMemcpyAsync(b, ctx->bathDev, sizeof b, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (bath) bath->chiT = b[B_REC_CHIT];
    if (ener) {
        ener->conint = b[B_REC_CONINT];
        ener->consv = b[B_REC_CONSV];
    }
    return MDG_OK;
}
