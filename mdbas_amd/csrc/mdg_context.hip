// mdg_context.hip -- the C ABI of libmdgpu (include/mdgpu.h): context,
// topology, the drop-in mirrors of the reference entry points and the
// device-resident MD loop.
#include "mdg_internal.h"

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>

#define TOLLIST_ 0.01            // global.h:59
#define SQRTPI_C 1.772453850905516
#define PI_C 3.141592653589793
#define TWOPI_C 6.283185307179586

// ---------------------------------------------------------------- helpers

// init.c:840-908, restated: reciprocal vectors divided by det, widths,
// volume.  The reference computes the same expressions, so the result is
// bit-identical to its init_box.
static void box_from_mirror(const MDG_PBC *in, MdgBox &b)
{
#pragma clang fp contract(off)
    b.type = in->type;
    b.a1 = in->a1; b.a2 = in->a2; b.a3 = in->a3;
    b.b1 = in->b1; b.b2 = in->b2; b.b3 = in->b3;
    b.c1 = in->c1; b.c2 = in->c2; b.c3 = in->c3;
    b.u1 = b.b2 * b.c3 - b.c2 * b.b3;
    b.u2 = b.c1 * b.b3 - b.b1 * b.c3;
    b.u3 = b.b1 * b.c2 - b.c1 * b.b2;
    b.v1 = b.c2 * b.a3 - b.a2 * b.c3;
    b.v2 = b.a1 * b.c3 - b.c1 * b.a3;
    b.v3 = b.c1 * b.a2 - b.a1 * b.c2;
    b.w1 = b.a2 * b.b3 - b.b2 * b.a3;
    b.w2 = b.b1 * b.a3 - b.a1 * b.b3;
    b.w3 = b.a1 * b.b2 - b.b1 * b.a2;
    b.det = b.a1 * b.u1 + b.a2 * b.u2 + b.a3 * b.u3;
    b.vol = fabs(b.det);
    b.pa = b.vol / sqrt(b.u1 * b.u1 + b.u2 * b.u2 + b.u3 * b.u3);
    b.pb = b.vol / sqrt(b.v1 * b.v1 + b.v2 * b.v2 + b.v3 * b.v3);
    b.pc = b.vol / sqrt(b.w1 * b.w1 + b.w2 * b.w2 + b.w3 * b.w3);
    b.u = 1.0 / b.pa;
    b.v = 1.0 / b.pb;
    b.w = 1.0 / b.pc;
    if (fabs(b.det) > 0.) {
        b.u1 /= b.det; b.u2 /= b.det; b.u3 /= b.det;
        b.v1 /= b.det; b.v2 /= b.det; b.v3 /= b.det;
        b.w1 /= b.det; b.w2 /= b.det; b.w3 /= b.det;
    } else {
        b.u1 = b.u2 = b.u3 = b.v1 = b.v2 = b.v3 = b.w1 = b.w2 = b.w3 = 0.;
    }
}

// io.c:714-723 derived constants plus the rest of the run settings.
static void params_from_mirror(const MDG_CTRL *ctrl, const MDG_PARAM *p, const MDG_EWALD *ew,
                               const MDG_NEIGH *ng, MdgParams &P)
{
#pragma clang fp contract(off)
    P.n = p->nAtom;
    P.cutOff = p->cutOff;
    P.cutOn = p->cutOn;
    P.delr = p->delr;
    P.chargeConst = p->chargeConst;
    P.alpha = ew ? ew->alpha : 0.;
    P.rcutOff = 1. / P.cutOff;
    P.cutOff2 = P.cutOff * P.cutOff;
    P.rcutOff2 = 1. / P.cutOff2;
    P.cutOn2 = P.cutOn * P.cutOn;
    double d = P.cutOff2 - P.cutOn2;
    P.switch2 = 1. / (d * d * d);
    double cutnb = P.cutOff + P.delr;        // list.c:906-907
    P.cutnb2 = cutnb * cutnb;
    P.timeStep = p->timeStep;
    P.rTimeStep = 1. / p->timeStep;
    P.tolShake = p->tolShake;
    P.maxCycle = p->maxCycle;
    P.update = ng ? ng->update : 20;
    P.linkRatio = ng ? std::min(std::max(ng->linkRatio, 1), 5) : 1;
    P.m1 = ew ? ew->m1max : 0;
    P.m2 = ew ? ew->m2max : 0;
    P.m3 = ew ? ew->m3max : 0;
    P.nbsp = ew ? ew->nbsp : 4;
    P.elecType = ctrl ? ctrl->elecType : 0;
    P.ewald = ctrl ? ctrl->keyEwald : 2;
    P.vdwOn = ctrl ? (ctrl->vdwType == 2) : 1;
}

// The same conversions for callers without a context.
void mdg_box_from(const MDG_PBC *in, MdgBox &b) { box_from_mirror(in, b); }

static void box_vectors(const MDG_PBC *in, double v[9])
{
    const double w[9] = {in->a1, in->a2, in->a3, in->b1, in->b2, in->b3, in->c1, in->c2, in->c3};
    for (int k = 0; k < 9; k++) v[k] = w[k];
}
void mdg_params_from(const MDG_CTRL *ctrl, const MDG_PARAM *p, MdgParams &P)
{
    params_from_mirror(ctrl, p, nullptr, nullptr, P);
}

template <typename T>
static int dmalloc(mdg_ctx *ctx, T **p, size_t n)
{
    HIPCHK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    return MDG_OK;
}

template <typename T>
static int h2d(mdg_ctx *ctx, T *d, const T *h, size_t n)
{
    if (n) HIPCHK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return MDG_OK;
}

template <typename T>
static int d2h(mdg_ctx *ctx, T *h, const T *d, size_t n)
{
    if (n) HIPCHK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    return MDG_OK;
}

static int sync(mdg_ctx *ctx)
{
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDG_OK;
}

int mdg_alloc_state(mdg_ctx *ctx)
{
    const int n = ctx->n;
    double **arr[] = {&ctx->x, &ctx->y, &ctx->z, &ctx->vx, &ctx->vy, &ctx->vz, &ctx->fx,
                      &ctx->fy, &ctx->fz, &ctx->q, &ctx->eps, &ctx->sig, &ctx->mass, &ctx->rmass,
                      &ctx->frx64, &ctx->fry64, &ctx->frz64};
    for (double **a : arr) RETCHK(dmalloc(ctx, a, n));
    RETCHK(dmalloc(ctx, &ctx->frx, n));
    RETCHK(dmalloc(ctx, &ctx->fry, n));
    RETCHK(dmalloc(ctx, &ctx->frz, n));
    RETCHK(dmalloc(ctx, &ctx->frozen, n));
    RETCHK(dmalloc(ctx, &ctx->nAtConst, n));
    RETCHK(dmalloc(ctx, &ctx->exOff, n + 1));
    RETCHK(dmalloc(ctx, &ctx->epart, PART_TOTAL));
    RETCHK(dmalloc(ctx, &ctx->scal, S_COUNT));
    RETCHK(dmalloc(ctx, &ctx->ctrl, CTRL_WORDS));
    RETCHK(mdg_fill_i32(ctx, ctx->ctrl, CTRL_WORDS, 0));
    RETCHK(mdg_fill_f64(ctx, ctx->scal, S_COUNT, 0.0));
    RETCHK(mdg_fill_i32(ctx, ctx->frozen, n, 0));
    RETCHK(mdg_fill_i32(ctx, ctx->nAtConst, n, 0));
    return MDG_OK;
}

// ---------------------------------------------------------------- exclusions

// list.c:108-894 (ordered, pair kept at the smaller index) and
// list.c:1125-1791 (symmetric).  Bonds and constraints are appended as
// given and recorded as connections; angle, dihedral and improper pairs
// only when absent, a new (a, d) of a dihedral also becoming a 1-4 pair;
// then the connectivity walk a-b-c(-d): (a, c) when absent, and for every
// d the reference's capacity check (a 1-4 path beyond 5 per dihedral is
// DIHE_NONPARAM_ERROR, 110) before (a, d) is added as exclusion and 1-4
// pair when absent.
int mdg_build_exclusions(mdg_ctx *ctx, int nBond, const MDG_BOND *bond, int nConst,
                         const MDG_CONSTRAINT *cons, int nAngle, const MDG_ANGLE *angle,
                         int nDihe, const MDG_DIHE *dihe, int nImpr, const MDG_DIHE *impr,
                         std::vector<int> *list14)
{
    const int n = ctx->n;
    const bool sym = ctx->keyLink != 0;
    auto bad = [n](int a) { return a < 0 || a >= n; };
    for (int i = 0; i < nBond; i++)
        if (bad(bond[i].a) || bad(bond[i].b)) return MDG_EARG;
    for (int i = 0; i < nConst; i++)
        if (bad(cons[i].a) || bad(cons[i].b)) return MDG_EARG;
    for (int i = 0; i < nAngle; i++)
        if (bad(angle[i].a) || bad(angle[i].b) || bad(angle[i].c)) return MDG_EARG;
    for (int i = 0; i < nDihe; i++)
        if (bad(dihe[i].a) || bad(dihe[i].b) || bad(dihe[i].c) || bad(dihe[i].d)) return MDG_EARG;
    for (int i = 0; i < nImpr; i++)
        if (bad(impr[i].a) || bad(impr[i].b) || bad(impr[i].c) || bad(impr[i].d)) return MDG_EARG;
    std::vector<std::vector<int>> ex(n), con(n);
    std::vector<int> l14;
    auto add = [&](int i, int j, bool dedupe) -> bool {
        int ii = i, jj = j;
        if (!sym && jj < ii) std::swap(ii, jj);
        if (dedupe)
            for (int v : ex[ii])
                if (v == jj) return false;
        ex[ii].push_back(jj);
        if (sym) ex[jj].push_back(ii);
        return true;
    };
    for (int i = 0; i < nBond; i++) {
        add(bond[i].a, bond[i].b, false);
        con[bond[i].a].push_back(bond[i].b);
        con[bond[i].b].push_back(bond[i].a);
    }
    for (int i = 0; i < nConst; i++) {
        add(cons[i].a, cons[i].b, false);
        con[cons[i].a].push_back(cons[i].b);
        con[cons[i].b].push_back(cons[i].a);
    }
    for (int i = 0; i < nAngle; i++) {
        add(angle[i].a, angle[i].b, true);
        add(angle[i].a, angle[i].c, true);
        add(angle[i].b, angle[i].c, true);
    }
    for (int i = 0; i < nDihe; i++) {
        const MDG_DIHE &d = dihe[i];
        add(d.a, d.b, true);
        add(d.a, d.c, true);
        if (add(d.a, d.d, true)) {
            l14.push_back(d.a);
            l14.push_back(d.d);
        }
        add(d.b, d.c, true);
        add(d.b, d.d, true);
        add(d.c, d.d, true);
    }
    for (int i = 0; i < nImpr; i++) {
        const MDG_DIHE &d = impr[i];
        add(d.a, d.b, true);
        add(d.a, d.c, true);
        add(d.a, d.d, true);
        add(d.b, d.c, true);
        add(d.b, d.d, true);
        add(d.c, d.d, true);
    }
    for (int ia = 0; ia < n; ia++)
        for (int ib : con[ia])
            for (int ic : con[ib]) {
                if (ic == ia) continue;
                add(ia, ic, true);
                for (int id : con[ic]) {
                    if (id == ia || id == ib) continue;
                    if ((long)l14.size() / 2 >= 5L * nDihe) {
                        ctx->err = "1-4 pair without dihedral parameters";
                        return 110;
                    }
                    if (add(ia, id, true)) {
                        l14.push_back(ia);
                        l14.push_back(id);
                    }
                }
            }
    if (!sym && n > 0 && !ex[n - 1].empty()) return 111;
    ctx->hExOff.assign(n + 1, 0);
    ctx->hExLst.clear();
    for (int i = 0; i < n; i++) {
        ctx->hExOff[i] = (int)ctx->hExLst.size();
        ctx->hExLst.insert(ctx->hExLst.end(), ex[i].begin(), ex[i].end());
    }
    ctx->hExOff[n] = (int)ctx->hExLst.size();
    ctx->exclBuilt = true;
    if (list14) *list14 = l14;
    return mdg_upload_exclusions(ctx);
}

int mdg_upload_exclusions(mdg_ctx *ctx)
{
    const int ne = (int)ctx->hExLst.size();
    if (ne > ctx->nExcl || !ctx->exLst) {
        if (ctx->exLst) HIPCHK(hipFree(ctx->exLst));
        RETCHK(dmalloc(ctx, &ctx->exLst, ne + 1));
        ctx->nExcl = ne;
    }
    RETCHK(h2d(ctx, ctx->exOff, ctx->hExOff.data(), ctx->n + 1));
    RETCHK(h2d(ctx, ctx->exLst, ctx->hExLst.data(), ne));
    ctx->exSpan = 0;
    for (int i = 0; i < ctx->n; i++)
        for (int k = ctx->hExOff[i]; k < ctx->hExOff[i + 1]; k++)
            ctx->exSpan = std::max(ctx->exSpan, std::abs(ctx->hExLst[k] - i));
    // Symmetric copy for the per-atom correction in the device-resident
    // step: every pair (i, j > i) of row i, with its multiplicity, is listed
    // under both atoms (the same pairs the reference's correction loop
    // visits, whichever orientation the table has).
    const int n = ctx->n;
    std::vector<int> cnt(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int k = ctx->hExOff[i]; k < ctx->hExOff[i + 1]; k++) {
            const int j = ctx->hExLst[k];
            if (j <= i) continue;
            cnt[i]++;
            cnt[j]++;
        }
    std::vector<int> off(n + 1, 0);
    for (int i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i];
    std::vector<int> lst(off[n] + 1), pos(off.begin(), off.end() - 1);
    for (int i = 0; i < n; i++)
        for (int k = ctx->hExOff[i]; k < ctx->hExOff[i + 1]; k++) {
            const int j = ctx->hExLst[k];
            if (j <= i) continue;
            lst[pos[i]++] = j;
            lst[pos[j]++] = i;
        }
    if (off[n] > ctx->nExSym || !ctx->exSymLst) {
        if (ctx->exSymLst) HIPCHK(hipFree(ctx->exSymLst));
        RETCHK(dmalloc(ctx, &ctx->exSymLst, off[n] + 1));
        ctx->nExSym = off[n];
    }
    if (!ctx->exSymOff) RETCHK(dmalloc(ctx, &ctx->exSymOff, n + 1));
    RETCHK(h2d(ctx, ctx->exSymOff, off.data(), n + 1));
    RETCHK(h2d(ctx, ctx->exSymLst, lst.data(), off[n]));
    return sync(ctx);
}

// Exclusion table from the reference's (exclList, exclPair) arguments.
static int exclusions_from_args(mdg_ctx *ctx, int **exclList, const int *exclPair)
{
    const int n = ctx->n;
    ctx->hExOff.assign(n + 1, 0);
    ctx->hExLst.clear();
    for (int i = 0; i < n; i++) {
        ctx->hExOff[i] = (int)ctx->hExLst.size();
        for (int k = 0; k < exclPair[i]; k++) ctx->hExLst.push_back(exclList[i][k]);
    }
    ctx->hExOff[n] = (int)ctx->hExLst.size();
    return mdg_upload_exclusions(ctx);
}

// ---------------------------------------------------------------- create

// Default list-pruning distance (mixed): on when every box edge leaves
// room for the prune kernel's cluster-relative test (half an edge above
// cutOff + skin plus a few angstroms of cluster extent), else off.
static double default_prune_skin(const mdg_ctx *ctx)
{
    if (ctx->precision != MDG_MIXED) return 0.;
    const double skin = std::min(MDG_PRUNE_SKIN, ctx->P.delr);
    auto roomy = [&](const MdgBox &B) {
        const double e = std::min(B.a1, B.type == 1 ? B.a1 : std::min(B.b2, B.c3));
        return 0.5 * e > ctx->P.cutOff + skin + 3.0;
    };
    if (ctx->bt) {
        for (const MdgBox &B : ctx->bt->hBox)
            if (!roomy(B)) return 0.;
        return skin;
    }
    return roomy(ctx->B) ? skin : 0.;
}

extern "C" mdg_ctx *mdg_create(const MDG_CTRL *ctrl, const MDG_PARAM *param, const MDG_PBC *box,
                               const MDG_EWALD *ewald, const MDG_NEIGH *neigh, int precision,
                               int *status)
{
    auto fail = [&](mdg_ctx *ctx, int code) -> mdg_ctx * {
        if (status) *status = code;
        if (ctx) {
            fprintf(stderr, "mdg_create: %s\n", ctx->err.c_str());
            mdg_destroy(ctx);
        }
        return nullptr;
    };
    if (!param || !box || param->nAtom <= 0) return fail(nullptr, MDG_EARG);
    mdg_ctx *ctx = new mdg_ctx();
    ctx->precision = precision == MDG_MIXED ? MDG_MIXED : MDG_FP64;
    ctx->n = param->nAtom;
    params_from_mirror(ctrl, param, ewald, neigh, ctx->P);
    box_from_mirror(box, ctx->B);
    box_vectors(box, ctx->boxKey);
    ctx->ensCreate = ctrl ? (int)ctrl->ens : 0;
    if (ctx->B.type < 1 || ctx->B.type > 3) {
        ctx->err = "a periodic box is required";
        return fail(ctx, MDG_EUNSUPPORTED);
    }
    if (ctx->precision == MDG_MIXED && ctx->B.type == 3) {
        // the mixed pair kernel's minimum image is per axis
        ctx->err = "mixed precision needs a cubic or orthorhombic box";
        return fail(ctx, MDG_EUNSUPPORTED);
    }
    if (hipGetDevice(&ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evFork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evJoin, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evRecFork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evRecDone, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evPruneDone, hipEventDisableTiming) != hipSuccess) {
        ctx->err = "no HIP device / stream";
        return fail(ctx, MDG_EHIP);
    }
    // makelist's first-call builder choice (list.c:38-57)
    double cutnb = ctx->P.cutOff + ctx->P.delr;
    int r = ctx->P.linkRatio;
    ctx->nlc[0] = (int)(ctx->B.pa * (double)r / cutnb);
    ctx->nlc[1] = (int)(ctx->B.pb * (double)r / cutnb);
    ctx->nlc[2] = (int)(ctx->B.pc * (double)r / cutnb);
    ctx->keyLink = 1;
    if (ctx->nlc[0] < 3 || ctx->nlc[1] < 3 || ctx->nlc[2] < 3) ctx->keyLink = 0;
    if (ctx->nlc[0] * ctx->nlc[1] * ctx->nlc[2] <= 27) ctx->keyLink = 0;
    if (ctrl && ctrl->noLink) ctx->keyLink = 0;
    if (ctx->keyLink && r != 1) {
        ctx->err = "link cells are implemented for linkRatio 1 (the 14-cell half shell)";
        return fail(ctx, MDG_EUNSUPPORTED);
    }
    if (ctx->keyLink && (ctx->nlc[0] > 1023 || ctx->nlc[1] > 1023 || ctx->nlc[2] > 1023)) {
        ctx->err = "more than 1023 link cells along an axis";
        return fail(ctx, MDG_EUNSUPPORTED);
    }
    ctx->ncell = ctx->nlc[0] * ctx->nlc[1] * ctx->nlc[2];
    if (int e = mdg_alloc_state(ctx)) return fail(ctx, e);
    ctx->m1 = ctx->P.m1;
    ctx->m2 = ctx->P.m2;
    ctx->m3 = ctx->P.m3;
    if (ctx->P.ewald != 0 && ctx->m1 > 0 && ctx->m2 > 0 && ctx->m3 > 0)
        if (int e = mdg_spme_setup(ctx)) return fail(ctx, e);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, MDG_EHIP);
    if (status) *status = MDG_OK;
    ctx->pruneSkin = default_prune_skin(ctx);
    return ctx;
}

// Replica batch: nrep independent boxes (cubic or orthorhombic, each
// small enough for the O(N^2) list of list.c:896-1040) with their own SPME
// meshes, concatenated atom by atom.  ctrl / param / ewald / neigh carry
// the settings all replicas share (param->nAtom is ignored).
extern "C" mdg_ctx *mdg_batch_create(const MDG_CTRL *ctrl, const MDG_PARAM *param,
                                     const MDG_EWALD *ewald, const MDG_NEIGH *neigh, int nrep,
                                     const int nAtom[], const MDG_PBC box[], const int mesh[],
                                     int precision, int *status)
{
    auto fail = [&](mdg_ctx *ctx, int code) -> mdg_ctx * {
        if (status) *status = code;
        if (ctx) {
            fprintf(stderr, "mdg_batch_create: %s\n", ctx->err.c_str());
            mdg_destroy(ctx);
        }
        return nullptr;
    };
    if (!param || !ewald || !box || !nAtom || !mesh || nrep <= 0) return fail(nullptr, MDG_EARG);
    mdg_ctx *ctx = new mdg_ctx();
    ctx->bt = new MdgBatch();
    MdgBatch &T = *ctx->bt;
    T.R = nrep;
    ctx->precision = precision == MDG_MIXED ? MDG_MIXED : MDG_FP64;
    params_from_mirror(ctrl, param, ewald, neigh, ctx->P);
    T.hAoff.assign(nrep + 1, 0);
    T.hBox.resize(nrep);
    T.hMesh.assign(mesh, mesh + nrep);
    T.hColOff.assign(nrep + 1, 0);
    T.hNfx.resize(nrep);
    T.hNfy.resize(nrep);
    const double cutnb = ctx->P.cutOff + ctx->P.delr;
    for (int r = 0; r < nrep; r++) {
        if (nAtom[r] <= 0 || mesh[r] < 4) {
            ctx->err = "every replica needs atoms and a mesh of at least 4";
            return fail(ctx, MDG_EARG);
        }
        T.hAoff[r + 1] = T.hAoff[r] + nAtom[r];
        box_from_mirror(&box[r], T.hBox[r]);
        const MdgBox &B = T.hBox[r];
        if (B.type != 1 && B.type != 2) {
            ctx->err = "batch replicas need cubic or orthorhombic boxes";
            return fail(ctx, MDG_EUNSUPPORTED);
        }
        if (ctx->P.ewald != 2) {
            ctx->err = "a replica batch runs the Ewald/SPME method (keyEwald 2)";
            return fail(ctx, MDG_EUNSUPPORTED);
        }
        // makelist's builder choice (list.c:38-57): this path is the O(N^2) one
        const int nx = (int)(B.pa / cutnb), ny = (int)(B.pb / cutnb), nz = (int)(B.pc / cutnb);
        const bool link = !(nx < 3 || ny < 3 || nz < 3 || nx * ny * nz <= 27 || (ctrl && ctrl->noLink));
        if (link) {
            ctx->err = "a replica large enough for link cells: create it as its own context";
            return fail(ctx, MDG_EUNSUPPORTED);
        }
        T.hNfx[r] = std::max(1, (int)(B.pa / 4.5));
        T.hNfy[r] = std::max(1, (int)(B.pb / 4.5));
        T.hColOff[r + 1] = T.hColOff[r] + T.hNfx[r] * T.hNfy[r];
    }
    ctx->n = T.hAoff[nrep];
    ctx->P.n = ctx->n;
    ctx->B = T.hBox[0];
    ctx->keyLink = 0;
    if (hipGetDevice(&ctx->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evFork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evJoin, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evRecFork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evRecDone, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->evPruneDone, hipEventDisableTiming) != hipSuccess) {
        ctx->err = "no HIP device / stream";
        return fail(ctx, MDG_EHIP);
    }
    if (int e = mdg_alloc_state(ctx)) return fail(ctx, e);
    std::vector<int> rep(ctx->n);
    for (int r = 0; r < nrep; r++)
        for (int i = T.hAoff[r]; i < T.hAoff[r + 1]; i++) rep[i] = r;
    int e = MDG_OK;
    auto up = [&](auto **d, const auto *h, size_t cnt) {
        if (e) return;
        if ((e = dmalloc(ctx, d, cnt))) return;
        e = h2d(ctx, *d, h, cnt);
    };
    up(&T.aoff, T.hAoff.data(), nrep + 1);
    up(&T.arep, rep.data(), ctx->n);
    up(&T.boxes, T.hBox.data(), nrep);
    up(&T.mesh, T.hMesh.data(), nrep);
    up(&T.colOff, T.hColOff.data(), nrep + 1);
    up(&T.nfx, T.hNfx.data(), nrep);
    up(&T.nfy, T.hNfy.data(), nrep);
    if (e) return fail(ctx, e);
    if ((e = dmalloc(ctx, &T.goff, nrep + 1)) || (e = dmalloc(ctx, &T.eSelf, nrep)) ||
        (e = dmalloc(ctx, &T.eNet, nrep)) || (e = dmalloc(ctx, &T.stepCtr, nrep)) ||
        (e = dmalloc(ctx, &T.masks, (size_t)nrep * 2 * 4)) ||   // [stage][replica][4 words]
        (e = dmalloc(ctx, &T.atomE, 2 * (size_t)ctx->n)))
        return fail(ctx, e);
    if ((e = mdg_b_spme_setup(ctx))) return fail(ctx, e);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, MDG_EHIP);
    if (status) *status = MDG_OK;
    ctx->pruneSkin = default_prune_skin(ctx);
    return ctx;
}

extern "C" int mdg_batch_size(const mdg_ctx *ctx) { return ctx && ctx->bt ? ctx->bt->R : 0; }

extern "C" void mdg_destroy(mdg_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (int k = 0; k < 2; k++)
        if (ctx->graphExec[k]) hipGraphExecDestroy(ctx->graphExec[k]);
    mdg_spme_free(ctx);
    mdg_bonded_free(ctx);
    void *ptrs[] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz, ctx->fx, ctx->fy, ctx->fz,
                    ctx->q, ctx->eps, ctx->sig, ctx->mass, ctx->rmass, ctx->frozen, ctx->nAtConst,
                    ctx->ca, ctx->cb, ctx->rc2, ctx->dd, ctx->grpAtomOff, ctx->grpAtoms,
                    ctx->grpConOff, ctx->grpCons, ctx->grpCycle, ctx->grpVir, ctx->ctrl,
                    ctx->exOff, ctx->exLst, ctx->xu, ctx->yu, ctx->zu, ctx->cellOf, ctx->cellCnt,
                    ctx->cellStart, ctx->lcKeys, ctx->lcS3, ctx->rl.neighList, ctx->rl.neighPair,
                    ctx->rl.neighOrder, ctx->rl.cnt, ctx->rl.off, ctx->cl.slotAtom,
                    ctx->cl.atomSlot, ctx->cl.bb, ctx->cl.bbh, ctx->cl.cpOff, ctx->cl.cpJ,
                    ctx->cl.cpMask, ctx->cl.cpI, ctx->cl.tmpCnt, ctx->cl.tmpJ, ctx->cl.tmpMask,
                    ctx->cl.colCnt, ctx->cl.colStart, ctx->cl.colPad, ctx->cl.colSlot,
                    ctx->cl.sortTmp, ctx->cl.sortKey, ctx->cl.cluCol, ctx->cl.pk, ctx->sx, ctx->sy, ctx->sz, ctx->sp32,
                    ctx->sq, ctx->seps, ctx->ssig, ctx->fdx, ctx->fdy, ctx->fdz, ctx->scal,
                    ctx->fix, ctx->frx, ctx->fry, ctx->frz, ctx->frx64,
                    ctx->fry64, ctx->frz64, ctx->epart, ctx->recDev, ctx->theta, ctx->theta32,
                    ctx->grid64, ctx->cgrid64, ctx->grid32, ctx->cgrid32, ctx->gridq,
                    ctx->fftWork, ctx->cl.chunk, ctx->cl.chOff, ctx->shkScr, ctx->shkVir, ctx->shkFirst, ctx->shkKend, ctx->shkRing,
                    ctx->shkMask, ctx->exSymOff, ctx->exSymLst, ctx->adjQ, ctx->adjE,
                    ctx->adjS, ctx->adjPot, ctx->binCnt, ctx->binBucket, ctx->binOut,
                    ctx->binNOut, ctx->splI, ctx->splB, ctx->splD,
                    ctx->cl.pJ, ctx->cl.pMask, ctx->cl.pCnt, ctx->dispPart, ctx->ensSave,
                    ctx->gridU, ctx->hvpBuf, ctx->bathDev, ctx->cl.cO, ctx->cl.cpD,
                    ctx->cl.pD, ctx->spr, ctx->relPart};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (ctx->bt) {
        MdgBatch &T = *ctx->bt;
        mdg_b_spme_free(ctx);
        void *bp[] = {T.aoff, T.goff, T.arep, T.grep, T.colOff, T.nfx, T.nfy, T.mesh, T.gridOff,
                      T.hOff, T.boxes, T.eSelf, T.eNet, T.cluStart, T.crep, T.chE, T.atomE,
                      T.grpKin, T.repInfl, T.repRec, T.masks, T.rec, T.stepCtr, T.binOff,
                      T.tileOff, T.tileRep};
        for (void *p : bp)
            if (p) hipFree(p);
        delete ctx->bt;
        ctx->bt = nullptr;
    }
    for (int k = 0; k < 8; k++)
        if (ctx->scanBuf[k]) hipFree(ctx->scanBuf[k]);
    for (hipEvent_t e : ctx->ev)
        if (e) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    if (ctx->stream2) hipStreamDestroy(ctx->stream2);
    if (ctx->evFork) hipEventDestroy(ctx->evFork);
    if (ctx->evJoin) hipEventDestroy(ctx->evJoin);
    if (ctx->evRecFork) hipEventDestroy(ctx->evRecFork);
    if (ctx->evRecDone) hipEventDestroy(ctx->evRecDone);
    if (ctx->evPruneDone) hipEventDestroy(ctx->evPruneDone);
    delete ctx;
}

extern "C" const char *mdg_last_error(const mdg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
extern "C" int mdg_precision(const mdg_ctx *ctx) { return ctx ? ctx->precision : -1; }
extern "C" int mdg_key_link(const mdg_ctx *ctx) { return ctx ? ctx->keyLink : -1; }

// spme.c:376-389 and 680 (self and net-charge energies), computed for
// every new set of charges.
static void charge_sums(mdg_ctx *ctx, const double q[])
{
    double systq = 0., q2 = 0.;
    for (int i = 0; i < ctx->n; i++) {
        systq += q[i];
        q2 += q[i] * q[i];
    }
    ctx->systq = systq;
    ctx->eSelf = -ctx->P.chargeConst * ctx->P.alpha * q2 / SQRTPI_C;
    ctx->eNet = ctx->B.vol > 0 && ctx->P.alpha > 0
        ? -(0.5 * PI_C * ctx->P.chargeConst) * ((systq / ctx->P.alpha) * (systq / ctx->P.alpha) / ctx->B.vol)
        : 0.;
}

// A box the caller changed between calls (the NPT integrators rescale it):
// the link-cell counts of the next list update (list.c:2380-2396), the
// net-charge energy (spme.c:680) and the influence function follow it.
int mdg_refresh_box(mdg_ctx *ctx, const MDG_PBC *box)
{
    if (!box || ctx->bt) return MDG_OK;
    double v[9];
    box_vectors(box, v);
    if (box->type == ctx->B.type && std::equal(v, v + 9, ctx->boxKey)) return MDG_OK;
    if (box->type != ctx->B.type) {
        ctx->err = "the box type changed";
        return MDG_EUNSUPPORTED;
    }
    if (ctx->precision == MDG_MIXED) {
        // the mixed kernels keep box-dependent state on the device path
        ctx->err = "a changing box needs an FP64 context";
        return MDG_EUNSUPPORTED;
    }
    box_from_mirror(box, ctx->B);
    std::copy(v, v + 9, ctx->boxKey);
    if (ctx->keyLink) {
        const double cutnb = ctx->P.cutOff + ctx->P.delr;
        const int r = ctx->P.linkRatio;
        ctx->nlc[0] = (int)(ctx->B.pa * (double)r / cutnb);
        ctx->nlc[1] = (int)(ctx->B.pb * (double)r / cutnb);
        ctx->nlc[2] = (int)(ctx->B.pc * (double)r / cutnb);
        if (ctx->nlc[0] < 2 * r + 1 || ctx->nlc[1] < 2 * r + 1 || ctx->nlc[2] < 2 * r + 1) {
            ctx->err = "the box shrank below three link cells along an axis";
            return 412;
        }
        ctx->ncell = ctx->nlc[0] * ctx->nlc[1] * ctx->nlc[2];
    }
    ctx->eNet = ctx->B.vol > 0 && ctx->P.alpha > 0
        ? -(0.5 * PI_C * ctx->P.chargeConst) *
              ((ctx->systq / ctx->P.alpha) * (ctx->systq / ctx->P.alpha) / ctx->B.vol)
        : 0.;
    if (ctx->theta) RETCHK(mdg_spme_rebox(ctx));
    mdg_invalidate_graphs(ctx);
    return sync(ctx);
}

// Per-replica self and net-charge energies (spme.c:376-389, 680 with each
// replica's own charges and volume).
static int batch_charge_sums(mdg_ctx *ctx, const double q[])
{
    MdgBatch &T = *ctx->bt;
    T.hESelf.assign(T.R, 0.);
    T.hENet.assign(T.R, 0.);
    T.hSystq.assign(T.R, 0.);
    for (int r = 0; r < T.R; r++) {
        double sq = 0., q2 = 0.;
        for (int i = T.hAoff[r]; i < T.hAoff[r + 1]; i++) {
            sq += q[i];
            q2 += q[i] * q[i];
        }
        const double a = ctx->P.alpha, v = T.hBox[r].vol;
        T.hSystq[r] = sq;
        T.hESelf[r] = -ctx->P.chargeConst * a * q2 / SQRTPI_C;
        T.hENet[r] = -(0.5 * PI_C * ctx->P.chargeConst) * ((sq / a) * (sq / a) / v);
    }
    RETCHK(h2d(ctx, T.eSelf, T.hESelf.data(), T.R));
    RETCHK(h2d(ctx, T.eNet, T.hENet.data(), T.R));
    return sync(ctx);
}

// Replica of every constraint group and each replica's group range (the
// groups follow the atom order, so a replica's groups are contiguous).
static int batch_topology(mdg_ctx *ctx, const double q[])
{
    MdgBatch &T = *ctx->bt;
    const int ng = ctx->nGrp;
    std::vector<int> off(ng + 1), atoms(ctx->n);
    RETCHK(d2h(ctx, off.data(), ctx->grpAtomOff, ng + 1));
    RETCHK(d2h(ctx, atoms.data(), ctx->grpAtoms, ctx->n));
    RETCHK(sync(ctx));
    std::vector<int> rep(ctx->n), grep(std::max(ng, 1));
    for (int r = 0; r < T.R; r++)
        for (int i = T.hAoff[r]; i < T.hAoff[r + 1]; i++) rep[i] = r;
    T.hGoff.assign(T.R + 1, ng);
    for (int g = ng - 1; g >= 0; g--) {
        const int r = rep[atoms[off[g]]];
        for (int k = off[g]; k < off[g + 1]; k++)
            if (rep[atoms[k]] != r) {
                ctx->err = "a constraint group spans two replicas";
                return MDG_EARG;
            }
        if (g > 0 && rep[atoms[off[g - 1]]] > r) {
            ctx->err = "constraint groups out of replica order";
            return MDG_EARG;
        }
        grep[g] = r;
        T.hGoff[r] = g;
    }
    for (int r = T.R - 1; r >= 0; r--)   // replicas without groups: empty ranges
        if (T.hGoff[r] > T.hGoff[r + 1]) T.hGoff[r] = T.hGoff[r + 1];
    if (T.grep) HIPCHK(hipFree(T.grep));
    if (T.grpKin) HIPCHK(hipFree(T.grpKin));
    RETCHK(dmalloc(ctx, &T.grep, std::max(ng, 1)));
    RETCHK(dmalloc(ctx, &T.grpKin, std::max(ng, 1)));
    RETCHK(h2d(ctx, T.grep, grep.data(), ng));
    RETCHK(h2d(ctx, T.goff, T.hGoff.data(), T.R + 1));
    if (!ctx->shakeW3) {
        ctx->err = "replica batches integrate rigid three-site molecules";
        return MDG_EUNSUPPORTED;
    }
    return batch_charge_sums(ctx, q);
}

// shake.c:243-258 tests 0.5 * |r^2 - d^2| / d < tol with d = sqrt(rc2)
// (x 0.5 is exact, so q = |r^2 - d^2| / d < 2 tol).  A correctly rounded
// division is monotone in the numerator, so the test holds exactly when the
// numerator is below the smallest a with fl(a / d) >= 2 tol, found here by
// bisection over the bit patterns of non-negative doubles; the device then
// compares without dividing.
static double shake_threshold(double rc2, double tol)
{
    const double d = std::sqrt(rc2), T = 2.0 * tol;
    if (!(T > 0.0) || !(d > 0.0)) return 0.0;   // never converged / degenerate
    auto val = [](uint64_t u) {
        double v;
        std::memcpy(&v, &u, sizeof v);
        return v;
    };
    uint64_t lo = 0, hi;   // fl(val(lo) / d) < T <= fl(val(hi) / d)
    {
        double h = T * d * 4.0;
        std::memcpy(&hi, &h, sizeof hi);
    }
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (val(mid) / d >= T) hi = mid;
        else lo = mid;
    }
    return val(hi);
}

extern "C" int mdg_set_topology(mdg_ctx *ctx, const double q[], const double eps[],
                                const double sig[], const double mass[], const double rmass[],
                                const int frozen[], const int nAtConst[],
                                const MDG_CONSTRAINT constList[], int nConst)
{
    if (!ctx || !q || !eps || !sig || !mass || !rmass) return MDG_EARG;
    const int n = ctx->n;
    RETCHK(h2d(ctx, ctx->q, q, n));
    RETCHK(h2d(ctx, ctx->eps, eps, n));
    RETCHK(h2d(ctx, ctx->sig, sig, n));
    RETCHK(h2d(ctx, ctx->mass, mass, n));
    RETCHK(h2d(ctx, ctx->rmass, rmass, n));
    ctx->hasFrozen = false;
    if (frozen) {
        RETCHK(h2d(ctx, ctx->frozen, frozen, n));
        for (int i = 0; i < n; i++) ctx->hasFrozen |= frozen[i] != 0;
    }
    if (nAtConst) RETCHK(h2d(ctx, ctx->nAtConst, nAtConst, n));
    charge_sums(ctx, q);
    // constraints
    if (ctx->ca) {
        hipFree(ctx->ca); hipFree(ctx->cb); hipFree(ctx->rc2); hipFree(ctx->dd);
        hipFree(ctx->shkThr);
    }
    ctx->nConst = nConst;
    ctx->hCons.assign(constList, constList + nConst);
    std::vector<int> a(std::max(nConst, 1)), b(std::max(nConst, 1));
    std::vector<double> r2(std::max(nConst, 1));
    for (int c = 0; c < nConst; c++) {
        a[c] = constList[c].a;
        b[c] = constList[c].b;
        r2[c] = constList[c].rc2;
    }
    RETCHK(dmalloc(ctx, &ctx->ca, nConst));
    RETCHK(dmalloc(ctx, &ctx->cb, nConst));
    RETCHK(dmalloc(ctx, &ctx->rc2, nConst));
    RETCHK(dmalloc(ctx, &ctx->dd, 3 * (size_t)nConst));
    RETCHK(h2d(ctx, ctx->ca, a.data(), nConst));
    RETCHK(h2d(ctx, ctx->cb, b.data(), nConst));
    RETCHK(h2d(ctx, ctx->rc2, r2.data(), nConst));
    {
        std::vector<double> th(std::max(nConst, 1));
        for (int c = 0; c < nConst; c++) th[c] = shake_threshold(r2[c], ctx->P.tolShake);
        RETCHK(dmalloc(ctx, &ctx->shkThr, nConst));
        RETCHK(h2d(ctx, ctx->shkThr, th.data(), nConst));
    }
    if (ctx->grpAtomOff) {
        hipFree(ctx->grpAtomOff); hipFree(ctx->grpAtoms); hipFree(ctx->grpConOff);
        hipFree(ctx->grpCons); hipFree(ctx->grpCycle); hipFree(ctx->grpVir);
    }
    RETCHK(mdg_build_groups(ctx, constList, nConst));
    RETCHK(mdg_build_exclusions(ctx, 0, nullptr, nConst, constList, 0, nullptr));
    if (ctx->bt) RETCHK(batch_topology(ctx, q));
    mdg_invalidate_graphs(ctx);
    return sync(ctx);
}

// New per-atom charges and LJ parameters for an existing topology (the
// parameters an adjoint differentiates); the device list's slot copies
// follow at once.
extern "C" int mdg_set_nonbond_params(mdg_ctx *ctx, const double q[], const double eps[],
                                      const double sig[])
{
    if (!ctx || !q || !eps || !sig) return MDG_EARG;
    const int n = ctx->n;
    RETCHK(h2d(ctx, ctx->q, q, n));
    RETCHK(h2d(ctx, ctx->eps, eps, n));
    RETCHK(h2d(ctx, ctx->sig, sig, n));
    charge_sums(ctx, q);
    if (ctx->bt) RETCHK(batch_charge_sums(ctx, q));
    mdg_invalidate_graphs(ctx);
    if (ctx->cl.cpJ && ctx->sx) RETCHK(mdg_slot_params(ctx));
    return sync(ctx);
}

// ---------------------------------------------------------------- drop-in: makelist

static int download_ref_list(mdg_ctx *ctx, MDG_NEIGH *neigh, int **neighList, int **neighPair,
                             int **neighOrder, bool first)
{
    const int n = ctx->n;
    RefList &R = ctx->rl;
    if (first) {
        *neighPair = (int *)malloc(n * sizeof(int));
        *neighOrder = (int *)malloc(n * sizeof(int));
        // list.c:2110 / 953 sizing
        neigh->sizeList = (int)((double)R.count * (1. + 2. * TOLLIST_)) + 1;
        *neighList = (int *)malloc((size_t)neigh->sizeList * sizeof(int));
    } else if (R.count > neigh->sizeList) {
        while (neigh->sizeList < R.count)
            neigh->sizeList = (int)((double)neigh->sizeList * (1. + TOLLIST_)) + 1;
        *neighList = (int *)realloc(*neighList, (size_t)neigh->sizeList * sizeof(int));
    }
    if (!*neighList || !*neighPair || !*neighOrder) return MDG_EARG;
    RETCHK(d2h(ctx, *neighList, R.neighList, R.count));
    RETCHK(d2h(ctx, *neighPair, R.neighPair, n));
    RETCHK(d2h(ctx, *neighOrder, R.neighOrder, n));
    return sync(ctx);
}

extern "C" int mdg_makelist(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_PBC *box,
                            MDG_NEIGH *neigh, MDG_CONSTRAINT constList[], MDG_BOND bond[],
                            MDG_ANGLE angle[], MDG_DIHE dihe[], MDG_DIHE impr[], double x[],
                            double y[], double z[], int frozen[], int **neighList,
                            int **neighPair, int **neighOrder, int **neighList14,
                            int ***exclList, int **exclPair)
{
    if (!ctx || !ctrl || !param || !neigh || !x || !y || !z) return MDG_EARG;
    if (box) RETCHK(mdg_refresh_box(ctx, box));
    const int n = ctx->n;
    if (param->nAtom != n) return MDG_EARG;
    if ((param->nBond > 0 && !bond) || (param->nAngle > 0 && !angle) ||
        (param->nDihedral > 0 && !dihe) || (param->nImproper > 0 && !impr))
        return MDG_EARG;
    if (ctrl->newjob == 1) {
        neigh->linkRatio = std::min(std::max(neigh->linkRatio, 1), 5);
        ctrl->keyLink = ctx->keyLink;
        if (frozen) {
            RETCHK(h2d(ctx, ctx->frozen, (const int *)frozen, n));
        }
        std::vector<int> l14;
        RETCHK(mdg_build_exclusions(ctx, param->nBond, bond, ctrl->keyConstH ? param->nConst : 0,
                                    constList, param->nAngle, angle, param->nDihedral, dihe,
                                    param->nImproper, impr, &l14));
        // exclusion table in the reference layout (N-1 rows ordered, N rows symmetric)
        const int rows = ctx->keyLink ? n : n - 1;
        *exclPair = (int *)malloc(n * sizeof(int));
        *exclList = (int **)malloc(std::max(rows, 1) * sizeof(int *));
        for (int i = 0; i < n; i++) (*exclPair)[i] = ctx->hExOff[i + 1] - ctx->hExOff[i];
        for (int i = 0; i < rows; i++) {
            (*exclList)[i] = (int *)malloc(std::max((*exclPair)[i], 1) * sizeof(int));
            for (int k = 0; k < (*exclPair)[i]; k++)
                (*exclList)[i][k] = ctx->hExLst[ctx->hExOff[i] + k];
        }
        // list.c:864-873: the 1-4 pairs as (i, j) couples
        neigh->nPair14 = (int)(l14.size() / 2);
        if (neighList14) {
            *neighList14 = (int *)malloc(std::max<size_t>(l14.size(), 1) * sizeof(int));
            std::copy(l14.begin(), l14.end(), *neighList14);
        }
        RETCHK(h2d(ctx, ctx->x, (const double *)x, n));
        RETCHK(h2d(ctx, ctx->y, (const double *)y, n));
        RETCHK(h2d(ctx, ctx->z, (const double *)z, n));
        RETCHK(mdg_build_ref_list(ctx, true));
        if (ctx->keyLink) {
            // list.c:1927: the first link-cell build wraps the coordinates
            RETCHK(d2h(ctx, x, ctx->x, n));
            RETCHK(d2h(ctx, y, ctx->y, n));
            RETCHK(d2h(ctx, z, ctx->z, n));
            neigh->nCells = ctx->ncell;
        }
        RETCHK(download_ref_list(ctx, neigh, neighList, neighPair, neighOrder, true));
        ctrl->newjob = 0;
    } else if (param->step % neigh->update == 0) {
        RETCHK(h2d(ctx, ctx->x, (const double *)x, n));
        RETCHK(h2d(ctx, ctx->y, (const double *)y, n));
        RETCHK(h2d(ctx, ctx->z, (const double *)z, n));
        RETCHK(mdg_build_ref_list(ctx, false));
        RETCHK(download_ref_list(ctx, neigh, neighList, neighPair, neighOrder, false));
    }
    int c[CTRL_WORDS];
    RETCHK(d2h(ctx, c, ctx->ctrl, CTRL_WORDS));
    RETCHK(sync(ctx));
    if (c[CTRL_SEG_OVF]) {
        ctx->err = "a link cell holds more atoms than one LDS sort segment";
        return MDG_EUNSUPPORTED;
    }
    return MDG_OK;
}

// ---------------------------------------------------------------- drop-in: energies

// Upload the list in the reference format with per-position offsets.
static int upload_ref_list(mdg_ctx *ctx, const int *neighList, const int *neighPair,
                           const int *neighOrder)
{
    const int n = ctx->n;
    RefList &R = ctx->rl;
    std::vector<long> off(n + 1, 0);
    for (int l = 0; l < n; l++) off[l + 1] = off[l] + neighPair[neighOrder[l]];
    const long total = off[n];
    if (!R.off) RETCHK(dmalloc(ctx, &R.off, n + 1));
    if (!R.neighOrder) RETCHK(dmalloc(ctx, &R.neighOrder, n));
    if (!R.neighPair) RETCHK(dmalloc(ctx, &R.neighPair, n));
    if (!R.cnt) RETCHK(dmalloc(ctx, &R.cnt, n));
    if (total > R.cap) {
        if (R.neighList) HIPCHK(hipFree(R.neighList));
        R.cap = total + 1024;
        RETCHK(dmalloc(ctx, &R.neighList, R.cap));
    }
    R.count = total;
    RETCHK(h2d(ctx, R.off, off.data(), n + 1));
    RETCHK(h2d(ctx, R.neighOrder, neighOrder, n));
    RETCHK(h2d(ctx, R.neighList, neighList, total));
    return sync(ctx);
}

static int upload_xyzq(mdg_ctx *ctx, const double *x, const double *y, const double *z,
                       const double *q, const double *eps, const double *sig)
{
    const int n = ctx->n;
    RETCHK(h2d(ctx, ctx->x, x, n));
    RETCHK(h2d(ctx, ctx->y, y, n));
    RETCHK(h2d(ctx, ctx->z, z, n));
    if (q) RETCHK(h2d(ctx, ctx->q, q, n));
    if (eps) RETCHK(h2d(ctx, ctx->eps, eps, n));
    if (sig) RETCHK(h2d(ctx, ctx->sig, sig, n));
    return MDG_OK;
}

static int refresh_charge_terms(mdg_ctx *ctx, const double *q)
{
    charge_sums(ctx, q);
    return MDG_OK;
}

extern "C" int mdg_spme_energy(mdg_ctx *ctx, MDG_PARAM *param, MDG_EWALD *ewald, MDG_PBC *box,
                               const double x[], const double y[], const double z[], double fx[],
                               double fy[], double fz[], const double q[], double stress[6],
                               double *virEwaldRec, double *eEwaldRec)
{
    (void)param; (void)ewald; (void)box;
    if (!ctx || !x || !fx || !q || !ctx->theta) return MDG_EARG;
    const int n = ctx->n;
    RETCHK(upload_xyzq(ctx, x, y, z, q, nullptr, nullptr));
    RETCHK(refresh_charge_terms(ctx, q));
    RETCHK(h2d(ctx, ctx->fx, (const double *)fx, n));
    RETCHK(h2d(ctx, ctx->fy, (const double *)fy, n));
    RETCHK(h2d(ctx, ctx->fz, (const double *)fz, n));
    double e[8];
    RETCHK(mdg_spme64(ctx, ctx->fx, ctx->fy, ctx->fz, e));
    RETCHK(d2h(ctx, fx, ctx->fx, n));
    RETCHK(d2h(ctx, fy, ctx->fy, n));
    RETCHK(d2h(ctx, fz, ctx->fz, n));
    RETCHK(sync(ctx));
    if (eEwaldRec) *eEwaldRec = e[0];
    if (virEwaldRec) *virEwaldRec = e[1];
    if (stress)
        for (int k = 0; k < 6; k++) stress[k] = e[2 + k];
    return MDG_OK;
}

// energy.c:450-643 on the device: uploads positions, parameters, list and
// exclusions and forces fx (the caller's +=), runs SPME (keyEwald 2), the
// real-space pairs and the excluded-pair correction into ctx->fx..; the
// forces stay on the device.  e[0..3] = elec, vdw, virelec, virvdw to add.
static int ewald_on_device(mdg_ctx *ctx, const MDG_CTRL *ctrl, const double *x, const double *y,
                           const double *z, const double *fx, const double *fy, const double *fz,
                           const double *q, const double *eps, const double *sig,
                           const int *neighList, const int *neighPair, const int *neighOrder,
                           int **exclList, const int *exclPair, double e[4])
{
    const int n = ctx->n;
    if (ctrl->keyEwald == 1) {
        ctx->err = "classic k-space Ewald (keyEwald 1) is not on this path";
        return MDG_EUNSUPPORTED;
    }
    ctx->P.vdwOn = ctrl->vdwType == 2;
    RETCHK(upload_xyzq(ctx, x, y, z, q, eps, sig));
    RETCHK(upload_ref_list(ctx, neighList, neighPair, neighOrder));
    RETCHK(exclusions_from_args(ctx, exclList, exclPair));
    RETCHK(h2d(ctx, ctx->fx, fx, n));
    RETCHK(h2d(ctx, ctx->fy, fy, n));
    RETCHK(h2d(ctx, ctx->fz, fz, n));
    double erec[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ctrl->keyEwald == 2) {
        if (!ctx->theta) return MDG_EARG;
        RETCHK(refresh_charge_terms(ctx, q));
        RETCHK(mdg_spme64(ctx, ctx->fx, ctx->fy, ctx->fz, erec));
    }
    RETCHK(mdg_pair_ref64(ctx, 0, ctx->fx, ctx->fy, ctx->fz, ctx->rl.neighList, ctx->rl.off,
                          ctx->rl.neighOrder, ctx->scal + S_REF));
    RETCHK(mdg_excl_corr64(ctx, ctx->fx, ctx->fy, ctx->fz, ctx->scal + S_REFCORR));
    double s[8];
    RETCHK(d2h(ctx, s, ctx->scal + S_REF, 8));
    RETCHK(sync(ctx));
    // energy.c:623-627
    e[0] = s[0] + erec[0] + s[6];
    e[1] = s[1];
    e[2] = s[2] + erec[1] + s[7];
    e[3] = s[3];
    return MDG_OK;
}

// energy.c:216-345 on the device (see ewald_on_device).
static int nonbond_on_device(mdg_ctx *ctx, const MDG_CTRL *ctrl, const double *x, const double *y,
                             const double *z, const double *fx, const double *fy,
                             const double *fz, const double *q, const double *eps,
                             const double *sig, const int *neighList, const int *neighPair,
                             const int *neighOrder, double e[4])
{
    const int n = ctx->n;
    // energy.c:45-65: FULL and NOELEC evaluate no electrostatics
    int et = ctrl->elecType;
    if (et == 0) et = 1;
    if (et < 1 || et > 4) return MDG_EUNSUPPORTED;
    ctx->P.vdwOn = ctrl->vdwType == 2;
    RETCHK(upload_xyzq(ctx, x, y, z, q, eps, sig));
    RETCHK(upload_ref_list(ctx, neighList, neighPair, neighOrder));
    RETCHK(h2d(ctx, ctx->fx, fx, n));
    RETCHK(h2d(ctx, ctx->fy, fy, n));
    RETCHK(h2d(ctx, ctx->fz, fz, n));
    RETCHK(mdg_pair_ref64(ctx, et, ctx->fx, ctx->fy, ctx->fz, ctx->rl.neighList, ctx->rl.off,
                          ctx->rl.neighOrder, ctx->scal + S_REF));
    double s[6];
    RETCHK(d2h(ctx, s, ctx->scal + S_REF, 6));
    RETCHK(sync(ctx));
    e[0] = s[0];
    e[1] = s[1];
    e[2] = s[2];
    e[3] = s[3];
    return MDG_OK;
}

static int download_forces(mdg_ctx *ctx, double *fx, double *fy, double *fz)
{
    RETCHK(d2h(ctx, fx, ctx->fx, ctx->n));
    RETCHK(d2h(ctx, fy, ctx->fy, ctx->n));
    RETCHK(d2h(ctx, fz, ctx->fz, ctx->n));
    return sync(ctx);
}

extern "C" int mdg_ewald_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                                MDG_EWALD *ewald, MDG_PBC *box, const double x[], const double y[],
                                const double z[], double fx[], double fy[], double fz[],
                                const double q[], const double eps[], const double sig[],
                                const int neighList[], const int neighPair[],
                                const int neighOrder[], int **exclList, const int exclPair[])
{
    (void)param; (void)ewald;
    if (!ctx || !ctrl || !ener || !x || !fx || !q || !neighPair || !neighOrder) return MDG_EARG;
    if (box) RETCHK(mdg_refresh_box(ctx, box));
    double e[4];
    RETCHK(ewald_on_device(ctx, ctrl, x, y, z, fx, fy, fz, q, eps, sig, neighList, neighPair,
                           neighOrder, exclList, exclPair, e));
    RETCHK(download_forces(ctx, fx, fy, fz));
    ener->elec += e[0];
    ener->vdw += e[1];
    ener->virelec += e[2];
    ener->virvdw += e[3];
    return MDG_OK;
}

extern "C" int mdg_nonbond_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param,
                                  MDG_ENERGY *ener, MDG_PBC *box, const double x[],
                                  const double y[], const double z[], double fx[], double fy[],
                                  double fz[], const double q[], const double eps[],
                                  const double sig[], const int neighList[],
                                  const int neighPair[], const int neighOrder[])
{
    (void)param;
    if (!ctx || !ctrl || !ener || !x || !fx || !q || !neighPair || !neighOrder) return MDG_EARG;
    if (box) RETCHK(mdg_refresh_box(ctx, box));
    double e[4];
    RETCHK(nonbond_on_device(ctx, ctrl, x, y, z, fx, fy, fz, q, eps, sig, neighList, neighPair,
                             neighOrder, e));
    RETCHK(download_forces(ctx, fx, fy, fz));
    ener->elec += e[0];
    ener->vdw += e[1];
    ener->virelec += e[2];
    ener->virvdw += e[3];
    return MDG_OK;
}

extern "C" int mdg_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                          MDG_EWALD *ewald, MDG_PBC *box, MDG_NEIGH *neigh, MDG_BOND bond[],
                          MDG_BOND ub[], MDG_ANGLE angle[], MDG_DIHE dihe[], MDG_DIHE impr[],
                          const double x[], const double y[], const double z[], double vx[],
                          double vy[], double vz[], double fx[], double fy[], double fz[],
                          const double q[], const double eps[], const double sig[],
                          const double eps14[], const double sig14[], const int frozen[],
                          const int neighList[], const int neighPair[], const int neighOrder[],
                          const int neighList14[], int **exclList, const int exclPair[])
{
    (void)ewald;
    if (!ctx || !ctrl || !param || !ener || !box || !x || !fx || !q || !neighPair || !neighOrder)
        return MDG_EARG;
    if (param->nAtom != ctx->n) return MDG_EARG;
    RETCHK(mdg_refresh_box(ctx, box));
    if ((param->nBond > 0 && !bond) || (param->nUb > 0 && !ub) || (param->nAngle > 0 && !angle) ||
        (param->nDihedral > 0 && !dihe) || (param->nImproper > 0 && !impr))
        return MDG_EARG;
    const int n14 = ctrl->keyNb14 && neigh ? neigh->nPair14 : 0;
    if (n14 > 0 && (!neighList14 || !eps14 || !sig14)) return MDG_EARG;
    // energy.c:115-144
    ener->elec = ener->vdw = ener->bond = ener->ub = ener->ang = ener->dihe = ener->impr = 0.;
    ener->virelec = ener->virvdw = ener->virbond = ener->virub = ener->virpot = 0.;
    box->stress1 = box->stress2 = box->stress3 = box->stress4 = box->stress5 = 0.;
    box->stress6 = box->stress7 = box->stress8 = box->stress9 = 0.;
    const int n = ctx->n;
    std::vector<double> zero(n, 0.);
    // energy.c:151-165: Ewald or plain nonbonded terms (into the zeroed
    // forces), each followed by its 1-4 pairs when keyNb14 is set
    double e[4];
    if (ctrl->keyEwald != 0)
        RETCHK(ewald_on_device(ctx, ctrl, x, y, z, zero.data(), zero.data(), zero.data(), q, eps,
                               sig, neighList, neighPair, neighOrder, exclList, exclPair, e));
    else
        RETCHK(nonbond_on_device(ctx, ctrl, x, y, z, zero.data(), zero.data(), zero.data(), q,
                                 eps, sig, neighList, neighPair, neighOrder, e));
    ener->elec += e[0];
    ener->vdw += e[1];
    ener->virelec += e[2];
    ener->virvdw += e[3];
    // energy.c:157-189: 1-4 pairs, bonds, angles, Urey-Bradley, dihedrals,
    // impropers, all in one device kernel
    MdgBonded &D = ctx->bd;
    D.scal14 = param->scal14;
    D.ewald14 = ctrl->keyEwald != 0;
    D.elec14 = ctrl->elecType;
    D.vdw14 = ctrl->vdwType == 2;
    RETCHK(mdg_bonded_upload(ctx, param->nBond, bond, param->nUb, ub, param->nAngle, angle,
                             param->nDihedral, dihe, param->nImproper, impr, n14, neighList14,
                             eps14, sig14));
    RETCHK(mdg_bonded_forces(ctx, ctx->fx, ctx->fy, ctx->fz, nullptr));
    double b[11];
    RETCHK(d2h(ctx, b, ctx->scal + S_BONDED, 11));
    RETCHK(download_forces(ctx, fx, fy, fz));
    ener->elec += b[7];
    ener->vdw += b[8];
    ener->virelec += b[9];
    ener->virvdw += b[10];
    ener->bond += b[0];
    ener->ub += b[1];
    ener->ang += b[2];
    ener->dihe += b[3];
    ener->impr += b[4];
    ener->virbond += b[5];
    ener->virub += b[6];
    // energy.c:193-196
    ener->pot = ener->elec + ener->vdw + ener->bond + ener->ang + ener->ub + ener->dihe + ener->impr;
    ener->virpot = ener->virbond + ener->virub + ener->virelec + ener->virvdw;
    if (param->nFrozen > 0 && frozen)   // utils.c:131-153
        for (int i = 0; i < n; i++)
            if (frozen[i]) vx[i] = vy[i] = vz[i] = fx[i] = fy[i] = fz[i] = 0.;
    return MDG_OK;
}

// ---------------------------------------------------------------- drop-in: integrator

extern "C" int mdg_vv_integrate(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                                MDG_PBC *box, MDG_BATH *bath, MDG_CONSTRAINT constList[],
                                double *x, double *y, double *z, double *vx, double *vy,
                                double *vz, double *fx, double *fy, double *fz, double *mass,
                                double *rmass, int *nAtConst, int stage)
{
    if (!ctx || !param || !ener || !x || !vx || !fx || !mass || !rmass) return MDG_EARG;
    const int ens = ctrl ? (int)ctrl->ens : 0;
    if (ens < 0 || ens > 4) return MDG_EARG;
    if (ens != 0 && (!bath || !box)) return MDG_EARG;
    if (box) RETCHK(mdg_refresh_box(ctx, box));
    const int n = ctx->n;
    ctx->P.timeStep = param->timeStep;
    ctx->P.rTimeStep = 1. / param->timeStep;
    ctx->P.tolShake = param->tolShake;
    ctx->P.maxCycle = param->maxCycle;
    if (!ctx->grpAtomOff) {
        if (nAtConst) RETCHK(h2d(ctx, ctx->nAtConst, (const int *)nAtConst, n));
        RETCHK(mdg_set_topology(ctx, std::vector<double>(n, 0.).data(),
                                std::vector<double>(n, 0.).data(),
                                std::vector<double>(n, 0.).data(), mass, rmass, nullptr, nAtConst,
                                constList, param->nConst));
    }
    RETCHK(h2d(ctx, ctx->mass, (const double *)mass, n));
    RETCHK(h2d(ctx, ctx->rmass, (const double *)rmass, n));
    if (nAtConst) RETCHK(h2d(ctx, ctx->nAtConst, (const int *)nAtConst, n));
    RETCHK(upload_xyzq(ctx, x, y, z, nullptr, nullptr, nullptr));
    RETCHK(h2d(ctx, ctx->vx, (const double *)vx, n));
    RETCHK(h2d(ctx, ctx->vy, (const double *)vy, n));
    RETCHK(h2d(ctx, ctx->vz, (const double *)vz, n));
    RETCHK(h2d(ctx, ctx->fx, (const double *)fx, n));
    RETCHK(h2d(ctx, ctx->fy, (const double *)fy, n));
    RETCHK(h2d(ctx, ctx->fz, (const double *)fz, n));
    double kin[8] = {0}, vs = 0.;
    if (ens != 0) {
        // integrate.c:1495-1506: the thermostat and barostat integrators
        RETCHK(mdg_vv_ensemble(ctx, param, ener, box, bath, ens, stage));
        RETCHK(d2h(ctx, x, ctx->x, n));
        RETCHK(d2h(ctx, y, ctx->y, n));
        RETCHK(d2h(ctx, z, ctx->z, n));
        RETCHK(d2h(ctx, vx, ctx->vx, n));
        RETCHK(d2h(ctx, vy, ctx->vy, n));
        RETCHK(d2h(ctx, vz, ctx->vz, n));
        return sync(ctx);
    }
    int st = mdg_vv_stage(ctx, stage, ctx->fx, ctx->fy, ctx->fz, kin, &vs, true);
    if (st) return st;
    RETCHK(d2h(ctx, x, ctx->x, n));
    RETCHK(d2h(ctx, y, ctx->y, n));
    RETCHK(d2h(ctx, z, ctx->z, n));
    RETCHK(d2h(ctx, vx, ctx->vx, n));
    RETCHK(d2h(ctx, vy, ctx->vy, n));
    RETCHK(d2h(ctx, vz, ctx->vz, n));
    RETCHK(sync(ctx));
    if (stage == 1) {
        if (param->nConst > 0) ener->virshake = vs;
    } else {
        ener->kin = kin[0];
        if (box) {   // integrate.c:1603-1611 (the SHAKE stress is zero in stage 2)
            box->stress1 += kin[1];
            box->stress2 += kin[2];
            box->stress3 += kin[3];
            box->stress4 += kin[2];
            box->stress5 += kin[4];
            box->stress6 += kin[5];
            box->stress7 += kin[3];
            box->stress8 += kin[5];
            box->stress9 += kin[6];
        }
    }
    return MDG_OK;
}
