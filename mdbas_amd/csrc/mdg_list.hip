// mdg_list.hip -- neighbour lists on the GPU.
//
// Two products:
//  * the reference-format list (neighList / neighPair / neighOrder) of
//    list.c, element for element: O(N^2) verlet_list (list.c:896-1123) or
//    link_cell_verlet_list (list.c:1793-2619), chosen as makelist does;
//  * a cluster-pair list for the device-resident kernels: atoms are sorted
//    into 8-atom clusters along spatial columns and every cluster pair
//    (I <= J) carries a 64-bit mask whose set bits are exactly the
//    reference list's pairs.  Masks are computed with the reference's own
//    FP64 predicate (no contraction), so the pair set is bit-exact.
//
// Cells and columns are sorted with a counting sort plus a per-segment
// bitonic sort in LDS on (key, atom) so every build is deterministic.
#include "mdg_internal.h"

// ---------------------------------------------------------------- common

__device__ inline unsigned int mdg_float_order(float f)
{
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline bool mdg_excluded(const int *__restrict__ exOff, const int *__restrict__ exLst,
                                    int i, int j)
{
    for (int k = exOff[i]; k < exOff[i + 1]; k++)
        if (exLst[k] == j) return true;
    return false;
}

// Block-aggregated counting: the atoms of one block that share a key are
// counted in LDS and added to the global counter with one atomic per key
// and block (thousands of atoms per cell or column would otherwise queue on
// a few global counters).  Every thread of the block must call these, with
// key = -1 when it has nothing to count.
#define MDG_AGG_KEYS 2048

__device__ inline void block_count(int key, int nkeys, int *cnt)
{
    __shared__ int h[MDG_AGG_KEYS];
    if (nkeys > MDG_AGG_KEYS) {
        if (key >= 0) atomicAdd(&cnt[key], 1);
        return;
    }
    for (int i = threadIdx.x; i < nkeys; i += blockDim.x) h[i] = 0;
    __syncthreads();
    if (key >= 0) atomicAdd(&h[key], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < nkeys; i += blockDim.x)
        if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// The thread's position among all atoms of its key (cursor[key] counts the
// positions handed out), in no particular order.
__device__ inline int block_rank(int key, int nkeys, int *cursor)
{
    __shared__ int h[MDG_AGG_KEYS], base[MDG_AGG_KEYS];
    if (nkeys > MDG_AGG_KEYS) return key >= 0 ? atomicAdd(&cursor[key], 1) : 0;
    for (int i = threadIdx.x; i < nkeys; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int r = key >= 0 ? atomicAdd(&h[key], 1) : 0;
    __syncthreads();
    for (int i = threadIdx.x; i < nkeys; i += blockDim.x)
        if (h[i]) base[i] = atomicAdd(&cursor[i], h[i]);
    __syncthreads();
    return key >= 0 ? base[key] + r : 0;
}

// utils.c:162-218 (image_update) on the device.
__global__ void k_image_update(MdgBox B, int n, double *x, double *y, double *z)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (B.type == 1) {
        x[i] = x[i] - B.a1 * mdg_nint(x[i] * B.u1);
        y[i] = y[i] - B.a1 * mdg_nint(y[i] * B.u1);
        z[i] = z[i] - B.a1 * mdg_nint(z[i] * B.u1);
    } else if (B.type == 2) {
        x[i] = x[i] - B.a1 * mdg_nint(x[i] * B.u1);
        y[i] = y[i] - B.b2 * mdg_nint(y[i] * B.v2);
        z[i] = z[i] - B.c3 * mdg_nint(z[i] * B.w3);
    } else if (B.type == 3) {
        double xt = x[i] * B.u1 + y[i] * B.u2 + z[i] * B.u3;
        double yt = x[i] * B.v1 + y[i] * B.v2 + z[i] * B.v3;
        double zt = x[i] * B.w1 + y[i] * B.w2 + z[i] * B.w3;
        xt -= mdg_nint(xt);
        yt -= mdg_nint(yt);
        zt -= mdg_nint(zt);
        x[i] = xt * B.a1 + yt * B.b1 + zt * B.c1;
        y[i] = xt * B.a2 + yt * B.b2 + zt * B.c2;
        z[i] = xt * B.a3 + yt * B.b3 + zt * B.c3;
    }
}

// Bitonic sort of each segment [start[s], start[s+1]) of 64-bit keys in
// LDS, ascending.  One block per segment.
__global__ void k_segment_sort(unsigned long long *__restrict__ keys, const int *__restrict__ start,
                               int *__restrict__ ctrl)
{
    __shared__ unsigned long long sk[MDG_SEG_MAX];
    const int s = blockIdx.x;
    const int b = start[s], len = start[s + 1] - b;
    if (len <= 1) return;
    if (len > MDG_SEG_MAX) {
        if (threadIdx.x == 0) atomicOr(&ctrl[CTRL_SEG_OVF], 1);
        return;
    }
    int p2 = 1;
    while (p2 < len) p2 <<= 1;
    for (int t = threadIdx.x; t < p2; t += blockDim.x) sk[t] = t < len ? keys[b + t] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < p2; t += blockDim.x) {
                int u = t ^ j;
                if (u > t) {
                    bool up = (t & k) == 0;
                    unsigned long long a = sk[t], c = sk[u];
                    if ((a > c) == up) { sk[t] = c; sk[u] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < len; t += blockDim.x) keys[b + t] = sk[t];
}

// ---------------------------------------------------------------- O(N^2)

// list.c:896-1040 count / fill passes, one thread per i (rows ascending).
template <bool FILL>
__global__ void k_verlet(MdgParams P, MdgBox B, const double *__restrict__ x,
                         const double *__restrict__ y, const double *__restrict__ z,
                         const int *__restrict__ frozen, const int *__restrict__ exOff,
                         const int *__restrict__ exLst, int *__restrict__ cnt,
                         const long *__restrict__ off, int *__restrict__ neighList)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    const int fi = frozen[i];
    int c = 0;
    long o = FILL ? off[i] : 0;
    for (int j = i + 1; j < P.n; j++) {
        double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
        double r2 = mdg_dist(B, dx, dy, dz);
        if (r2 <= P.cutnb2) {
            bool ex = (fi * frozen[j]) ? true : mdg_excluded(exOff, exLst, i, j);
            if (!ex) {
                if (FILL) neighList[o + c] = j;
                c++;
            }
        }
    }
    if (!FILL) cnt[i] = c;
}

__global__ void k_identity(int n, int *order, const int *cnt, int *neighPair)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        order[i] = i;
        neighPair[i] = cnt[i];
    }
}

// ---------------------------------------------------------------- link cells

// list.c:1932-1967: scaled coordinates and cell index (truncation, MIN).
__global__ void k_lc_assign(MdgParams P, MdgBox B, int nx, int ny, int nz,
                            const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, double *__restrict__ xu,
                            double *__restrict__ yu, double *__restrict__ zu,
                            int *__restrict__ cellOf, int *__restrict__ cellCnt)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int cell = -1;
    if (i < P.n) {
        double a = (x[i] * B.u1 + y[i] * B.u2 + z[i] * B.u3) + 0.5;
        double b = (x[i] * B.v1 + y[i] * B.v2 + z[i] * B.v3) + 0.5;
        double c = (x[i] * B.w1 + y[i] * B.w2 + z[i] * B.w3) + 0.5;
        xu[i] = a;
        yu[i] = b;
        zu[i] = c;
        int ix = (int)((double)nx * a), iy = (int)((double)ny * b), iz = (int)((double)nz * c);
        ix = ix <= nx - 1 ? ix : nx - 1;
        iy = iy <= ny - 1 ? iy : ny - 1;
        iz = iz <= nz - 1 ? iz : nz - 1;
        cell = ix + nx * (iy + ny * iz);
        cellOf[i] = cell;
    }
    // the counts serve the reference-format list; the cluster list needs
    // only the cells
    if (cellCnt) block_count(cell, nx * ny * nz, cellCnt);
}

// Place (cell, descending atom) keys into the cell segments; the segment
// sort then orders them, giving head/link chain order (list.c:1958-1980).
__global__ void k_lc_place(int n, int ncell, const int *__restrict__ cellOf,
                           const int *__restrict__ cellStart, int *__restrict__ cursor,
                           unsigned long long *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = i < n ? cellOf[i] : -1;
    const int r = block_rank(c, ncell, cursor);
    if (c >= 0) keys[cellStart[c] + r] = (unsigned long long)(0x7fffffff - i);
}

__global__ void k_lc_order(int n, const unsigned long long *__restrict__ keys, int *__restrict__ order,
                           const double *__restrict__ xu, const double *__restrict__ yu,
                           const double *__restrict__ zu, double *__restrict__ sxu,
                           double *__restrict__ syu, double *__restrict__ szu)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int i = 0x7fffffff - (int)(keys[p] & 0xffffffffull);
    order[p] = i;
    sxu[p] = xu[i];
    syu[p] = yu[i];
    szu[p] = zu[i];
}

__constant__ int c_tx[14] = {0, 1, 0, 0, -1, 1, 0, -1, 1, 0, -1, 1, -1, 1};
__constant__ int c_ty[14] = {0, 0, 1, 0, 1, 1, -1, 0, 0, 1, -1, -1, 1, 1};
__constant__ int c_tz[14] = {0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1};

// list.c:2130-2261: for position p (atom i in cell k), the 14 half-shell
// cells in table order; j runs over the cell's chain, or over the chain
// after i in i's own cell.  Count pass and fill pass.
template <bool FILL>
__global__ void k_lc_pairs(MdgParams P, MdgBox B, int nx, int ny, int nz,
                           const int *__restrict__ order, const int *__restrict__ cellOf,
                           const int *__restrict__ cellStart, const double *__restrict__ sxu,
                           const double *__restrict__ syu, const double *__restrict__ szu,
                           const int *__restrict__ frozen, const int *__restrict__ exOff,
                           const int *__restrict__ exLst, int *__restrict__ cnt,
                           const long *__restrict__ off, int *__restrict__ neighList,
                           int *__restrict__ neighPair)
{
#pragma clang fp contract(off)
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int i = order[p];
    const int k = cellOf[i];
    const int ix = k % nx, iy = (k / nx) % ny, iz = k / (nx * ny);
    const double xi = sxu[p], yi = syu[p], zi = szu[p];
    const int fi = frozen[i];
    int c = 0;
    long o = FILL ? off[p] : 0;
    for (int l = 0; l < 14; l++) {
        double cx = 0., cy = 0., cz = 0.;
        int jx = ix + c_tx[l], jy = iy + c_ty[l], jz = iz + c_tz[l];
        if (jx > nx - 1) { jx -= nx; cx = 1.; } else if (jx < 0) { jx += nx; cx = -1.; }
        if (jy > ny - 1) { jy -= ny; cy = 1.; } else if (jy < 0) { jy += ny; cy = -1.; }
        if (jz > nz - 1) { jz -= nz; cz = 1.; } else if (jz < 0) { jz += nz; cz = -1.; }
        int ll = jx + nx * (jy + ny * jz);
        int q0 = (ll == k) ? p + 1 : cellStart[ll];
        int q1 = cellStart[ll + 1];
        for (int q = q0; q < q1; q++) {
            double xt = sxu[q] - xi + cx;
            double yt = syu[q] - yi + cy;
            double zt = szu[q] - zi + cz;
            double xd = xt * B.a1 + yt * B.b1 + zt * B.c1;
            double yd = xt * B.a2 + yt * B.b2 + zt * B.c2;
            double zd = xt * B.a3 + yt * B.b3 + zt * B.c3;
            double r2 = xd * xd + yd * yd + zd * zd;
            if (r2 <= P.cutnb2) {
                int j = order[q];
                bool ex = (fi * frozen[j]) ? true : mdg_excluded(exOff, exLst, i, j);
                if (!ex) {
                    if (FILL) neighList[o + c] = j;
                    c++;
                }
            }
        }
    }
    if (!FILL) cnt[p] = c;
    else neighPair[i] = c;
}

static int ensure_int(mdg_ctx *ctx, int **p, size_t n)
{
    if (!*p) HIPCHK(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(int)));
    return MDG_OK;
}

static int ensure_double(mdg_ctx *ctx, double **p, size_t n)
{
    if (!*p) HIPCHK(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double)));
    return MDG_OK;
}

// Link-cell scaled coordinates, cells and the chain order; also used by the
// cluster list (for the link-cell pair predicate).
static int lc_sort(mdg_ctx *ctx, unsigned long long **keys, double **sxu, double **syu,
                   double **szu)
{
    const int n = ctx->n, nx = ctx->nlc[0], ny = ctx->nlc[1], nz = ctx->nlc[2];
    RETCHK(ensure_double(ctx, &ctx->xu, n));
    RETCHK(ensure_double(ctx, &ctx->yu, n));
    RETCHK(ensure_double(ctx, &ctx->zu, n));
    RETCHK(ensure_int(ctx, &ctx->cellOf, n));
    RETCHK(ensure_int(ctx, &ctx->cellCnt, 2 * (ctx->ncell + 1)));
    RETCHK(ensure_int(ctx, &ctx->cellStart, ctx->ncell + 1));
    RETCHK(mdg_fill_i32(ctx, ctx->cellCnt, 2 * (ctx->ncell + 1), 0));
    hipLaunchKernelGGL(k_lc_assign, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P, ctx->B,
                       nx, ny, nz, ctx->x, ctx->y, ctx->z, ctx->xu, ctx->yu, ctx->zu, ctx->cellOf,
                       ctx->cellCnt);
    LAUNCHCHK();
    RETCHK(mdg_scan_int(ctx, ctx->cellCnt, ctx->cellStart, ctx->ncell));
    int *cursor = ctx->cellCnt + ctx->ncell + 1;
    hipLaunchKernelGGL(k_lc_place, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, ctx->ncell,
                       ctx->cellOf, ctx->cellStart, cursor, *keys);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_segment_sort, dim3(ctx->ncell), dim3(256), 0, ctx->stream, *keys,
                       ctx->cellStart, ctx->ctrl);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_lc_order, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, *keys,
                       ctx->rl.neighOrder, ctx->xu, ctx->yu, ctx->zu, *sxu, *syu, *szu);
    LAUNCHCHK();
    return MDG_OK;
}

// Builds the reference-format list into ctx->rl (device) from ctx->x,y,z.
// wrap = the first-call link-cell build, which wraps x,y,z first.
int mdg_build_ref_list(mdg_ctx *ctx, bool wrap)
{
    const int n = ctx->n;
    RefList &R = ctx->rl;
    RETCHK(ensure_int(ctx, &R.neighPair, n));
    RETCHK(ensure_int(ctx, &R.neighOrder, n));
    RETCHK(ensure_int(ctx, &R.cnt, n));
    if (!R.off) HIPCHK(hipMalloc(&R.off, (n + 1) * sizeof(long)));
    if (ctx->keyLink) {
        if (wrap) {
            hipLaunchKernelGGL(k_image_update, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                               ctx->B, n, ctx->x, ctx->y, ctx->z);
            LAUNCHCHK();
        }
        if (!ctx->lcKeys) HIPCHK(hipMalloc(&ctx->lcKeys, n * sizeof(unsigned long long)));
        if (!ctx->lcS3) HIPCHK(hipMalloc(&ctx->lcS3, 3 * (size_t)n * sizeof(double)));
        unsigned long long *keys = ctx->lcKeys;
        double *s3 = ctx->lcS3;
        double *sxu = s3, *syu = s3 + n, *szu = s3 + 2 * n;
        RETCHK(lc_sort(ctx, &keys, &sxu, &syu, &szu));
        const int nx = ctx->nlc[0], ny = ctx->nlc[1], nz = ctx->nlc[2];
        hipLaunchKernelGGL((k_lc_pairs<false>), dim3(nblk(n, 128)), dim3(128), 0, ctx->stream,
                           ctx->P, ctx->B, nx, ny, nz, R.neighOrder, ctx->cellOf, ctx->cellStart,
                           sxu, syu, szu, ctx->frozen, ctx->exOff, ctx->exLst, R.cnt, nullptr,
                           nullptr, nullptr);
        LAUNCHCHK();
        RETCHK(mdg_scan_int_to_long(ctx, R.cnt, R.off, n));
        long total = 0;
        HIPCHK(hipMemcpyAsync(&total, R.off + n, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (total > R.cap) {
            if (R.neighList) HIPCHK(hipFree(R.neighList));
            R.cap = (long)(total * 1.2) + 1024;
            HIPCHK(hipMalloc(&R.neighList, R.cap * sizeof(int)));
        }
        R.count = total;
        hipLaunchKernelGGL((k_lc_pairs<true>), dim3(nblk(n, 128)), dim3(128), 0, ctx->stream,
                           ctx->P, ctx->B, nx, ny, nz, R.neighOrder, ctx->cellOf, ctx->cellStart,
                           sxu, syu, szu, ctx->frozen, ctx->exOff, ctx->exLst, R.cnt, R.off,
                           R.neighList, R.neighPair);
        LAUNCHCHK();
    } else {
        hipLaunchKernelGGL((k_verlet<false>), dim3(nblk(n, 128)), dim3(128), 0, ctx->stream,
                           ctx->P, ctx->B, ctx->x, ctx->y, ctx->z, ctx->frozen, ctx->exOff,
                           ctx->exLst, R.cnt, nullptr, nullptr);
        LAUNCHCHK();
        RETCHK(mdg_scan_int_to_long(ctx, R.cnt, R.off, n));
        long total = 0;
        HIPCHK(hipMemcpyAsync(&total, R.off + n, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (total > R.cap) {
            if (R.neighList) HIPCHK(hipFree(R.neighList));
            R.cap = (long)(total * 1.2) + 1024;
            HIPCHK(hipMalloc(&R.neighList, R.cap * sizeof(int)));
        }
        R.count = total;
        hipLaunchKernelGGL((k_verlet<true>), dim3(nblk(n, 128)), dim3(128), 0, ctx->stream,
                           ctx->P, ctx->B, ctx->x, ctx->y, ctx->z, ctx->frozen, ctx->exOff,
                           ctx->exLst, nullptr, R.off, R.neighList);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_identity, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n,
                           R.neighOrder, R.cnt, R.neighPair);
        LAUNCHCHK();
    }
    return MDG_OK;
}

// ---------------------------------------------------------------- clusters

// Column of each atom (xy grid in scaled coordinates) and its sort key
// (scaled z, then atom index).
__global__ void k_col_assign(MdgBox B, int n, int ncx, int ncy, const double *__restrict__ x,
                             const double *__restrict__ y, const double *__restrict__ z,
                             int *__restrict__ colOf, int *__restrict__ colCnt,
                             float *__restrict__ zkey)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int col = -1;
    if (i < n) {
        double a = x[i] * B.u1 + y[i] * B.u2 + z[i] * B.u3 + 0.5;
        double b = x[i] * B.v1 + y[i] * B.v2 + z[i] * B.v3 + 0.5;
        double c = x[i] * B.w1 + y[i] * B.w2 + z[i] * B.w3 + 0.5;
        a -= floor(a);
        b -= floor(b);
        c -= floor(c);
        int cx = min((int)(a * ncx), ncx - 1), cy = min((int)(b * ncy), ncy - 1);
        col = cx + ncx * cy;
        colOf[i] = col;
        zkey[i] = (float)c;
    }
    block_count(col, ncx * ncy, colCnt);
}

__global__ void k_col_place(int n, int ncol, const int *__restrict__ colOf,
                            const int *__restrict__ colStart, int *__restrict__ cursor,
                            const float *__restrict__ zkey, unsigned long long *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = i < n ? colOf[i] : -1;
    const int r = block_rank(c, ncol, cursor);
    if (c >= 0)
        keys[colStart[c] + r] = ((unsigned long long)mdg_float_order(zkey[i]) << 32) | (unsigned int)i;
}

__global__ void k_col_pad(int ncol, const int *__restrict__ colCnt, int *__restrict__ colPad)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol) colPad[c] = (colCnt[c] + MDG_CL - 1) / MDG_CL * MDG_CL;
}

// Slots of column c: [colSlot[c], colSlot[c] + colPad[c]); atoms fill the
// first colCnt[c] slots in z order, the rest are padding (-1).
__global__ void k_col_slots(int ncol, const int *__restrict__ colStart,
                            const int *__restrict__ colCnt, const int *__restrict__ colSlot,
                            const unsigned long long *__restrict__ keys, int *__restrict__ slotAtom,
                            int *__restrict__ atomSlot, int *__restrict__ cluCol)
{
    int c = blockIdx.x;
    int s0 = colSlot[c], np = colSlot[c + 1] - s0, na = colCnt[c], a0 = colStart[c];
    for (int t = threadIdx.x; t < np; t += blockDim.x) {
        int atom = -1;
        if (t < na) {
            atom = (int)(keys[a0 + t] & 0xffffffffull);
            atomSlot[atom] = s0 + t;
        }
        slotAtom[s0 + t] = atom;
        if ((t & (MDG_CL - 1)) == 0) cluCol[(s0 + t) / MDG_CL] = c;   // the cluster's column
    }
}

// Bounding box of each cluster from wrapped coordinates (centre, half
// extent, both FP32, extents rounded outward).
//
// With cO given (rectangular boxes), also the cluster's origin for the
// relative coordinates of the mixed pair kernel: its first atom plus the
// mean minimum-image offset of its atoms from that atom (a centre that does
// not depend on where the box's seams cut the cluster), fixed until the
// next build.
__global__ void k_cluster_bbox(MdgBox B, int nclu, const int *__restrict__ slotAtom,
                               const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, float4 *__restrict__ bb,
                               float4 *__restrict__ bbh, double *__restrict__ cO)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nclu) return;
    if (cO) {
        const double L[3] = {B.a1, B.type == 1 ? B.a1 : B.b2, B.type == 1 ? B.a1 : B.c3};
        const double iL[3] = {B.u1, B.type == 1 ? B.u1 : B.v2, B.type == 1 ? B.u1 : B.w3};
        double o[3] = {0., 0., 0.}, m[3] = {0., 0., 0.};
        int cnt = 0;
        for (int k = 0; k < MDG_CL; k++) {
            const int a = slotAtom[I * MDG_CL + k];
            if (a < 0) continue;
            const double p[3] = {x[a], y[a], z[a]};
            for (int d = 0; d < 3; d++) {
                if (cnt == 0) o[d] = p[d];
                const double u = p[d] - o[d];
                m[d] += u - L[d] * rint(u * iL[d]);
            }
            cnt++;
        }
        for (int d = 0; d < 3; d++) cO[(long)d * nclu + I] = cnt ? o[d] + m[d] / cnt : 0.;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < MDG_CL; k++) {
        int a = slotAtom[I * MDG_CL + k];
        if (a < 0) continue;
        // wrap into [-L/2, L/2) by scaled coordinates
        double s[3] = {x[a] * B.u1 + y[a] * B.u2 + z[a] * B.u3, x[a] * B.v1 + y[a] * B.v2 + z[a] * B.v3,
                       x[a] * B.w1 + y[a] * B.w2 + z[a] * B.w3};
        for (int d = 0; d < 3; d++) s[d] -= floor(s[d] + 0.5);
        double p[3] = {s[0] * B.a1 + s[1] * B.b1 + s[2] * B.c1, s[0] * B.a2 + s[1] * B.b2 + s[2] * B.c2,
                       s[0] * B.a3 + s[1] * B.b3 + s[2] * B.c3};
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
    }
    if (lo[0] > hi[0]) {   // empty cluster (cannot happen: columns are padded per 8)
        bb[I] = make_float4(0.f, 0.f, 0.f, 0.f);
        bbh[I] = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
        return;
    }
    bb[I] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                        (float)(0.5 * (lo[2] + hi[2])), 0.f);
    bbh[I] = make_float4((float)(0.5 * (hi[0] - lo[0])) + 1e-3f, (float)(0.5 * (hi[1] - lo[1])) + 1e-3f,
                         (float)(0.5 * (hi[2] - lo[2])) + 1e-3f, 0.f);
}

// Exact reference predicate for the unordered pair (i, j): the same FP64
// expression list.c evaluates for whichever builder the run uses, plus the
// frozen and exclusion tests.  The value does not depend on the order of
// i and j (negation is exact, nint and the cell shift are odd).
__device__ inline bool ref_pair_pred(const MdgParams &P, const MdgBox &B, int keyLink, int nx,
                                     int ny, int nz, int i, int j, const double *__restrict__ x,
                                     const double *__restrict__ y, const double *__restrict__ z,
                                     const double *__restrict__ xu, const double *__restrict__ yu,
                                     const double *__restrict__ zu, const int *__restrict__ cellOf,
                                     const int *__restrict__ frozen, const int *__restrict__ exOff,
                                     const int *__restrict__ exLst)
{
#pragma clang fp contract(off)
    double r2;
    if (keyLink) {
        int ci = cellOf[i], cj = cellOf[j];
        int ix = ci % nx, iy = (ci / nx) % ny, iz = ci / (nx * ny);
        int jx = cj % nx, jy = (cj / nx) % ny, jz = cj / (nx * ny);
        int d[3] = {jx - ix, jy - iy, jz - iz};
        int nn[3] = {nx, ny, nz};
        double c[3];
        for (int k = 0; k < 3; k++) {
            if (d[k] == 0 || d[k] == 1 || d[k] == -1) c[k] = 0.;
            else if (d[k] == -(nn[k] - 1)) c[k] = 1.;     // j's cell wrapped past the top
            else if (d[k] == nn[k] - 1) c[k] = -1.;       // j's cell wrapped past the bottom
            else return false;                             // not neighbouring cells
        }
        double xt = xu[j] - xu[i] + c[0];
        double yt = yu[j] - yu[i] + c[1];
        double zt = zu[j] - zu[i] + c[2];
        double xd = xt * B.a1 + yt * B.b1 + zt * B.c1;
        double yd = xt * B.a2 + yt * B.b2 + zt * B.c2;
        double zd = xt * B.a3 + yt * B.b3 + zt * B.c3;
        r2 = xd * xd + yd * yd + zd * zd;
    } else {
        int lo = min(i, j), hi = max(i, j);
        double dx = x[hi] - x[lo], dy = y[hi] - y[lo], dz = z[hi] - z[lo];
        r2 = mdg_dist(B, dx, dy, dz);
    }
    if (!(r2 <= P.cutnb2)) return false;
    if (frozen[i] * frozen[j]) return false;
    // the O(N^2) table keeps a pair at its smaller index only (list.c:882-890);
    // the link-cell table keeps it at both (list.c:1779-1787)
    return !mdg_excluded(exOff, exLst, min(i, j), max(i, j));
}

// ref_pair_pred on slot-ordered copies of the predicate inputs (the same
// FP64 expressions): atoms ia, ja with predicate coordinates (scaled for
// link cells, else positions) and link cells ci, cj.  The exclusion table
// is searched only when |ia - ja| is within the largest excluded span, and
// frozen flags only when the system has frozen atoms.
__device__ inline bool ref_pair_pred_slot(const MdgParams &P, const MdgBox &B, int keyLink, int nx,
                                          int ny, int nz, int ia, int ja, double xi, double yi,
                                          double zi, int ci, double xj, double yj, double zj,
                                          int cj, bool hasFrozen, const int *__restrict__ frozen,
                                          const int *__restrict__ exOff,
                                          const int *__restrict__ exLst, int exSpan)
{
#pragma clang fp contract(off)
    double r2;
    if (keyLink) {
        // ci, cj: cell coordinates packed 10 bits each
        const int ix = ci & 1023, iy = (ci >> 10) & 1023, iz = ci >> 20;
        const int jx = cj & 1023, jy = (cj >> 10) & 1023, jz = cj >> 20;
        int d[3] = {jx - ix, jy - iy, jz - iz};
        int nn[3] = {nx, ny, nz};
        double c[3];
        for (int k = 0; k < 3; k++) {
            if (d[k] == 0 || d[k] == 1 || d[k] == -1) c[k] = 0.;
            else if (d[k] == -(nn[k] - 1)) c[k] = 1.;     // j's cell wrapped past the top
            else if (d[k] == nn[k] - 1) c[k] = -1.;       // j's cell wrapped past the bottom
            else return false;                             // not neighbouring cells
        }
        double xt = xj - xi + c[0];
        double yt = yj - yi + c[1];
        double zt = zj - zi + c[2];
        double xd, yd, zd;
        if (B.type == 3) {
            xd = xt * B.a1 + yt * B.b1 + zt * B.c1;
            yd = xt * B.a2 + yt * B.b2 + zt * B.c2;
            zd = xt * B.a3 + yt * B.b3 + zt * B.c3;
        } else {
            // the off-diagonal terms are products with +-0: the sums equal
            // the diagonal products up to the sign of a zero, which r2 drops
            xd = xt * B.a1;
            yd = yt * B.b2;
            zd = zt * B.c3;
        }
        r2 = xd * xd + yd * yd + zd * zd;
    } else {
        // x[hi] - x[lo], as the reference orders the pair
        double dx, dy, dz;
        if (ia < ja) {
            dx = xj - xi;
            dy = yj - yi;
            dz = zj - zi;
        } else {
            dx = xi - xj;
            dy = yi - yj;
            dz = zi - zj;
        }
        r2 = mdg_dist(B, dx, dy, dz);
    }
    if (!(r2 <= P.cutnb2)) return false;
    if (hasFrozen && frozen[ia] * frozen[ja]) return false;
    const int lo = min(ia, ja), hi = max(ia, ja);
    return !(hi - lo <= exSpan && mdg_excluded(exOff, exLst, lo, hi));
}

// Per slot, in one 32-byte record: the predicate inputs of its atom (x, y,
// z) and, in the fourth double's halves, the link cell as packed
// coordinates (low) and the atom (high; -1 for padding, never tested).
__global__ void k_pred_slots(int nslot, int keyLink, int nx, int ny, const int *__restrict__ slotAtom,
                             const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const double *__restrict__ xu,
                             const double *__restrict__ yu, const double *__restrict__ zu,
                             const int *__restrict__ cellOf, double4 *__restrict__ pk)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int a = slotAtom[s];
    if (a < 0) {
        pk[s] = make_double4(0., 0., 0., __hiloint2double(-1, 0));
        return;
    }
    int cp = 0;
    if (keyLink) {
        // cell coordinates, packed (cells per axis stay below 1024)
        const int c = cellOf[a];
        cp = (c % nx) | (((c / nx) % ny) << 10) | ((c / (nx * ny)) << 20);
    }
    pk[s] = make_double4(keyLink ? xu[a] : x[a], keyLink ? yu[a] : y[a], keyLink ? zu[a] : z[a],
                         __hiloint2double(a, cp));
}

// One wave per I cluster: bbox test against every J >= I, then the 64
// exact atom-pair predicates of each surviving (I, J) as one ballot.
//
// Candidates are the clusters of the column rows around I's column: a
// cluster lies inside its column, so clusters more than reach = rlist / w
// + 1 columns apart (w the column width) cannot hold a pair within rlist.
// The columns of one row that are in reach form at most two contiguous
// cluster ranges (two when they wrap around the box).
#define CAND_BATCH 4             // candidate j clusters whose slot records load together

struct ColGrid {
    const int *colSlot;          // first slot of each column (ncol + 1)
    const int *cluCol;           // column of each cluster (-1 unused)
    int ncx, ncy, rx, ry;        // columns per row / rows, reach in each
};

__global__ void __launch_bounds__(256)
k_cluster_pairs(MdgParams P, MdgBox B, int keyLink, int nx, int ny, int nz, int nclu,
                const float4 *__restrict__ bb,
                const float4 *__restrict__ bbh, const double4 *__restrict__ pk,
                bool hasFrozen, const int *__restrict__ frozen,
                const int *__restrict__ exOff, const int *__restrict__ exLst, int exSpan, int kmax,
                int *__restrict__ tmpCnt, int *__restrict__ tmpJ,
                unsigned long long *__restrict__ tmpMask, int *__restrict__ ctrl, ColGrid CG)
{
    const int lane = threadIdx.x & 63;
    const int I = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (I >= nclu) return;
    const int colI = CG.cluCol[I];
    if (colI < 0) {
        if (lane == 0) tmpCnt[I] = 0;
        return;
    }
    const float4 ci = bb[I], hi = bbh[I];
    const float L0 = (float)B.a1, L1 = (float)B.b2, L2 = (float)B.c3;
    const float rl = (float)(P.cutOff + P.delr) + 0.01f;
    const int a = lane & 7, b = lane >> 3;   // i, j atom: mask bit = lane = 8 j + i
    const int si = I * MDG_CL + a;
    const double4 pi = pk[si];
    const int iat = __double2hiint(pi.w), cli = __double2loint(pi.w);
    const double xi = pi.x, yi = pi.y, zi = pi.z;
    const int cx = colI % CG.ncx, cy = colI / CG.ncx;
    // rows: all of them once when the reach covers the box, else cy +- ry
    const bool allRows = 2 * CG.ry + 1 >= CG.ncy, allCols = 2 * CG.rx + 1 >= CG.ncx;
    const int nrow = allRows ? CG.ncy : 2 * CG.ry + 1;
    int found = 0;
    for (int rr = 0; rr < nrow; rr++) {
        const int row = allRows ? rr : (cy - CG.ry + rr + CG.ncy) % CG.ncy;
        // the row's columns in reach as up to two index ranges [lo, hi)
        int rlo[2], rhi[2], nr = 1;
        if (allCols) {
            rlo[0] = 0;
            rhi[0] = CG.ncx;
        } else {
            const int lo = cx - CG.rx, hc = cx + CG.rx + 1;
            if (lo < 0) {
                rlo[0] = lo + CG.ncx; rhi[0] = CG.ncx; rlo[1] = 0; rhi[1] = hc; nr = 2;
            } else if (hc > CG.ncx) {
                rlo[0] = lo; rhi[0] = CG.ncx; rlo[1] = 0; rhi[1] = hc - CG.ncx; nr = 2;
            } else {
                rlo[0] = lo; rhi[0] = hc;
            }
        }
        for (int q = 0; q < nr; q++) {
            const int j0 = max(I, CG.colSlot[row * CG.ncx + rlo[q]] / MDG_CL);
            const int j1 = CG.colSlot[row * CG.ncx + rhi[q]] / MDG_CL;
            for (int J0 = j0; J0 < j1; J0 += 64) {
                const int J = J0 + lane;
                bool cand = false;
                if (J < j1) {
                    const float4 cj = bb[J], hj = bbh[J];
                    float dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
                    if (B.type == 1 || B.type == 2) {
                        dx -= L0 * rintf(dx / L0);
                        dy -= L1 * rintf(dy / L1);
                        dz -= L2 * rintf(dz / L2);
                    }
                    dx = fmaxf(fabsf(dx) - hi.x - hj.x, 0.f);
                    dy = fmaxf(fabsf(dy) - hi.y - hj.y, 0.f);
                    dz = fmaxf(fabsf(dz) - hi.z - hj.z, 0.f);
                    cand = B.type == 3 || (dx * dx + dy * dy + dz * dz <= rl * rl);
                }
                unsigned long long cm = __ballot(cand);
                // candidates in order, CAND_BATCH at a time: their slot
                // records are all requested before the first is tested
                while (cm) {
                    int src[CAND_BATCH];
                    int cnt = 0;
#pragma unroll
                    for (int t = 0; t < CAND_BATCH; t++) {
                        src[t] = cm ? __ffsll((long long)cm) - 1 : 0;
                        if (cm) {
                            cm &= cm - 1;
                            cnt++;
                        }
                    }
                    double4 pj[CAND_BATCH];
#pragma unroll
                    for (int t = 0; t < CAND_BATCH; t++)
                        pj[t] = pk[(J0 + (t < cnt ? src[t] : src[0])) * MDG_CL + b];
#pragma unroll
                    for (int t = 0; t < CAND_BATCH; t++) {
                        if (t >= cnt) break;
                        const int Jc = J0 + src[t];
                        const int jat = __double2hiint(pj[t].w), clj = __double2loint(pj[t].w);
                        bool pr = false;
                        if (iat >= 0 && jat >= 0 && (Jc != I || a < b))
                            pr = ref_pair_pred_slot(P, B, keyLink, nx, ny, nz, iat, jat, xi, yi,
                                                    zi, cli, pj[t].x, pj[t].y, pj[t].z, clj,
                                                    hasFrozen, frozen, exOff, exLst, exSpan);
                        const unsigned long long m = __ballot(pr);
                        if (m) {
                            if (found < kmax && lane == 0) {
                                tmpJ[(long)I * kmax + found] = Jc;
                                tmpMask[(long)I * kmax + found] = m;
                            }
                            found++;
                        }
                    }
                }
            }
        }
    }
    if (lane == 0) {
        tmpCnt[I] = min(found, kmax);
        if (found > kmax) atomicOr(&ctrl[CTRL_LIST_OVF], 1);
    }
}

// With cpD given: each entry's minimum image of O_J - O_I (cluster
// origins, rectangular box), rounded to FP32.
__global__ void k_cluster_compact(int nclu, int kmax, const int *__restrict__ tmpCnt,
                                  const int *__restrict__ tmpJ,
                                  const unsigned long long *__restrict__ tmpMask,
                                  const int *__restrict__ cpOff, long cap, int *__restrict__ cpJ,
                                  unsigned long long *__restrict__ cpMask, int *__restrict__ cpI,
                                  int *__restrict__ ctrl, MdgBox B, const double *__restrict__ cO,
                                  float4 *__restrict__ cpD)
{
    int I = blockIdx.x;
    int o = cpOff[I], c = tmpCnt[I];
    for (int t = threadIdx.x; t < c; t += blockDim.x) {
        if (o + t < cap) {
            const int J = tmpJ[(long)I * kmax + t];
            cpJ[o + t] = J;
            cpMask[o + t] = tmpMask[(long)I * kmax + t];
            cpI[o + t] = I;
            if (cpD) {
                const double L[3] = {B.a1, B.type == 1 ? B.a1 : B.b2, B.type == 1 ? B.a1 : B.c3};
                const double iL[3] = {B.u1, B.type == 1 ? B.u1 : B.v2, B.type == 1 ? B.u1 : B.w3};
                float v[3];
                for (int d = 0; d < 3; d++) {
                    const double u = cO[(long)d * nclu + J] - cO[(long)d * nclu + I];
                    v[d] = (float)(u - L[d] * rint(u * iL[d]));
                }
                cpD[o + t] = make_float4(v[0], v[1], v[2], 0.f);
            }
        } else if (t == 0) {
            atomicOr(&ctrl[CTRL_LIST_OVF], 2);
        }
    }
}

__global__ void k_chunk_count(int nclu, const int *__restrict__ tmpCnt, int *__restrict__ nch)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I < nclu) nch[I] = (tmpCnt[I] + MDG_CHUNK - 1) / MDG_CHUNK;
}

__global__ void k_chunk_fill(int nclu, const int *__restrict__ cpOff, const int *__restrict__ chOff,
                             int maxChunks, int2 *__restrict__ chunk, int *__restrict__ ctrl)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nclu) return;
    for (int c = chOff[I]; c < chOff[I + 1]; c++) {
        if (c < maxChunks) chunk[c] = make_int2(I, cpOff[I] + (c - chOff[I]) * MDG_CHUNK);
        else atomicOr(&ctrl[CTRL_LIST_OVF], 4);
    }
}

// Column sort shared by both builders: atoms placed by column (colOf,
// counts in C.colCnt), sorted by (z key, atom) inside each column, columns
// padded to whole clusters, slot <-> atom maps written.
static int sort_columns(mdg_ctx *ctx, const int *colOf, unsigned long long *keys)
{
    const int n = ctx->n;
    ClusterList &C = ctx->cl;
    RETCHK(mdg_scan_int(ctx, C.colCnt, C.colStart, C.ncol));
    hipLaunchKernelGGL(k_col_place, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, C.ncol,
                       colOf, C.colStart, C.colCnt + C.ncol + 1, C.sortKey, keys);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_segment_sort, dim3(C.ncol), dim3(256), 0, ctx->stream, keys, C.colStart,
                       ctx->ctrl);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_col_pad, dim3(nblk(C.ncol, 256)), dim3(256), 0, ctx->stream, C.ncol,
                       C.colCnt, C.colPad);
    LAUNCHCHK();
    RETCHK(mdg_scan_int(ctx, C.colPad, C.colSlot, C.ncol));
    RETCHK(mdg_fill_i32(ctx, C.cluCol, C.nclu, -1));
    hipLaunchKernelGGL(k_col_slots, dim3(C.ncol), dim3(64), 0, ctx->stream, C.ncol, C.colStart,
                       C.colCnt, C.colSlot, keys, C.slotAtom, C.atomSlot, C.cluCol);
    LAUNCHCHK();
    return MDG_OK;
}

// Pair-count scan, list compaction and work chunks, shared by both
// builders (C.tmpCnt / tmpJ / tmpMask filled by the pair search).  The
// first call sizes the list from its own count with headroom.
static int compact_cluster_list(mdg_ctx *ctx)
{
    ClusterList &C = ctx->cl;
    const int nclu = C.nclu;
    RETCHK(mdg_scan_int(ctx, C.tmpCnt, C.cpOff, nclu));
    if (C.cap == 0) {
        // first build: size the list from this count with headroom
        int total = 0;
        HIPCHK(hipMemcpyAsync(&total, C.cpOff + nclu, sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        C.cap = (long)(total * 1.5) + 4096;
        HIPCHK(hipMalloc(&C.cpJ, C.cap * sizeof(int)));
        HIPCHK(hipMalloc(&C.cpI, C.cap * sizeof(int)));
        HIPCHK(hipMalloc(&C.cpMask, C.cap * sizeof(unsigned long long)));
        C.maxChunks = (int)(C.cap / MDG_CHUNK) + nclu + 64;
        HIPCHK(hipMalloc(&C.chunk, (size_t)C.maxChunks * sizeof(int2)));
        HIPCHK(hipMalloc(&C.chOff, (nclu + 1) * sizeof(int)));
        if (ctx->precision == MDG_MIXED) {
            HIPCHK(hipMalloc(&C.pJ, C.cap * sizeof(int)));
            HIPCHK(hipMalloc(&C.pMask, C.cap * sizeof(unsigned long long)));
            HIPCHK(hipMalloc(&C.pCnt, (size_t)C.maxChunks * sizeof(int)));
            if (C.cO) {
                HIPCHK(hipMalloc(&C.cpD, C.cap * sizeof(float4)));
                HIPCHK(hipMalloc(&C.pD, C.cap * sizeof(float4)));
            }
        }
    }
    // the pruned list belongs to the previous list: the next slot update
    // starts over from the full one
    ctx->pruneFresh = true;
    hipLaunchKernelGGL(k_cluster_compact, dim3(nclu), dim3(64), 0, ctx->stream, nclu, C.kmax,
                       C.tmpCnt, C.tmpJ, C.tmpMask, C.cpOff, C.cap, C.cpJ, C.cpMask, C.cpI,
                       ctx->ctrl, ctx->B, C.cO, C.cpD);
    LAUNCHCHK();
    // work chunks of at most MDG_CHUNK entries for the pair kernel; the
    // per-I chunk counts reuse tmpCnt's space after the compaction
    int *nch = C.sortTmp;   // n ints, free again after the column sort (n >= nclu)
    hipLaunchKernelGGL(k_chunk_count, dim3(nblk(nclu, 256)), dim3(256), 0, ctx->stream, nclu,
                       C.tmpCnt, nch);
    LAUNCHCHK();
    RETCHK(mdg_scan_int(ctx, nch, C.chOff, nclu));
    hipLaunchKernelGGL(k_chunk_fill, dim3(nblk(nclu, 256)), dim3(256), 0, ctx->stream, nclu,
                       C.cpOff, C.chOff, C.maxChunks, C.chunk, ctx->ctrl);
    LAUNCHCHK();
    return MDG_OK;
}

// Builds the cluster-pair list from ctx->x,y,z (device).  Every call is
// graph-capturable once the buffers exist (first call sizes them).
int mdg_build_cluster_list(mdg_ctx *ctx)
{
    const int n = ctx->n;
    ClusterList &C = ctx->cl;
    if (!C.colCnt) {
        // xy columns of ~4.5 A: ~8 atoms per 4 A of column height in water
        double w = 4.5;
        C.nfx = std::max(1, (int)(ctx->B.pa / w));
        C.nfy = std::max(1, (int)(ctx->B.pb / w));
        C.ncol = C.nfx * C.nfy;
        int maxslot = n + (MDG_CL - 1) * C.ncol;
        int maxclu = maxslot / MDG_CL + 1;
        HIPCHK(hipMalloc(&C.colCnt, 2 * (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colStart, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colPad, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colSlot, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.sortTmp, n * sizeof(int)));
        HIPCHK(hipMalloc(&C.sortKey, n * sizeof(float)));
        HIPCHK(hipMalloc(&C.slotAtom, ((size_t)maxclu * MDG_CL + 64) * sizeof(int)));
        HIPCHK(hipMalloc(&C.atomSlot, n * sizeof(int)));
        HIPCHK(hipMalloc(&C.bb, maxclu * sizeof(float4)));
        HIPCHK(hipMalloc(&C.bbh, maxclu * sizeof(float4)));
        if (ctx->precision == MDG_MIXED && ctx->B.type != 3 && !ctx->bt)
            HIPCHK(hipMalloc(&C.cO, 3 * (size_t)maxclu * sizeof(double)));
        HIPCHK(hipMalloc(&C.cpOff, (maxclu + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.tmpCnt, maxclu * sizeof(int)));
        HIPCHK(hipMalloc(&C.cluCol, maxclu * sizeof(int)));
        const size_t ns = (size_t)maxclu * MDG_CL + 64;
        HIPCHK(hipMalloc(&C.pk, ns * sizeof(double4)));
        // neighbouring clusters within rc+skin plus two cluster diameters
        double rl = ctx->P.cutOff + ctx->P.delr + 9.0;
        double dens = (double)n / ctx->B.vol / MDG_CL;
        C.kmax = std::max(64, (int)(4.18879 * rl * rl * rl * dens * 0.75) + 64);
        C.kmax = std::min(C.kmax, maxclu);
        HIPCHK(hipMalloc(&C.tmpJ, (size_t)maxclu * C.kmax * sizeof(int)));
        HIPCHK(hipMalloc(&C.tmpMask, (size_t)maxclu * C.kmax * sizeof(unsigned long long)));
        C.nclu = maxclu;
        C.cap = 0;
        RETCHK(mdg_fill_i32(ctx, C.slotAtom, (long)maxclu * MDG_CL + 64, -1));
    }
    unsigned long long *keys = (unsigned long long *)C.tmpMask;   // reused as sort scratch
    int *colOf = C.sortTmp;
    RETCHK(mdg_fill_i32(ctx, C.slotAtom, (long)C.nclu * MDG_CL + 64, -1));
    RETCHK(mdg_fill_i32(ctx, C.colCnt, 2 * (C.ncol + 1), 0));
    hipLaunchKernelGGL(k_col_assign, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->B, n,
                       C.nfx, C.nfy, ctx->x, ctx->y, ctx->z, colOf, C.colCnt, C.sortKey);
    LAUNCHCHK();
    RETCHK(sort_columns(ctx, colOf, keys));
    // The number of clusters varies with the padding; kernels run over the
    // maximum and empty clusters produce nothing.
    const int nclu = C.nclu;
    hipLaunchKernelGGL(k_cluster_bbox, dim3(nblk(nclu, 128)), dim3(128), 0, ctx->stream, ctx->B,
                       nclu, C.slotAtom, ctx->x, ctx->y, ctx->z, C.bb, C.bbh, C.cO);
    LAUNCHCHK();
    if (ctx->keyLink) {
        // the link-cell predicate needs scaled coordinates and cells
        const int nx = ctx->nlc[0], ny = ctx->nlc[1], nz = ctx->nlc[2];
        RETCHK(ensure_double(ctx, &ctx->xu, n));
        RETCHK(ensure_double(ctx, &ctx->yu, n));
        RETCHK(ensure_double(ctx, &ctx->zu, n));
        RETCHK(ensure_int(ctx, &ctx->cellOf, n));
        hipLaunchKernelGGL(k_lc_assign, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P,
                           ctx->B, nx, ny, nz, ctx->x, ctx->y, ctx->z, ctx->xu, ctx->yu, ctx->zu,
                           ctx->cellOf, nullptr);
        LAUNCHCHK();
    }
    hipLaunchKernelGGL(k_pred_slots, dim3(nblk((long)nclu * MDG_CL, 256)), dim3(256), 0,
                       ctx->stream, nclu * MDG_CL, ctx->keyLink, ctx->nlc[0], ctx->nlc[1],
                       C.slotAtom, ctx->x, ctx->y, ctx->z,
                       ctx->xu, ctx->yu, ctx->zu, ctx->cellOf, C.pk);
    LAUNCHCHK();
    ColGrid CG;
    CG.colSlot = C.colSlot;
    CG.cluCol = C.cluCol;
    CG.ncx = C.nfx;
    CG.ncy = C.nfy;
    {
        const double rl = ctx->P.cutOff + ctx->P.delr;
        CG.rx = (int)(rl / (ctx->B.pa / C.nfx)) + 1;
        CG.ry = (int)(rl / (ctx->B.pb / C.nfy)) + 1;
    }
    hipLaunchKernelGGL(k_cluster_pairs, dim3(nblk(nclu, 4)), dim3(256), 0, ctx->stream, ctx->P,
                       ctx->B, ctx->keyLink, ctx->nlc[0], ctx->nlc[1], ctx->nlc[2], nclu,
                       C.bb, C.bbh, C.pk, ctx->hasFrozen,
                       ctx->frozen, ctx->exOff, ctx->exLst, ctx->exSpan, C.kmax, C.tmpCnt, C.tmpJ,
                       C.tmpMask, ctx->ctrl, CG);
    LAUNCHCHK();
    return compact_cluster_list(ctx);
}

// ---------------------------------------------------------------- replica batch

// Column of each atom in its own replica's xy grid (replica r owns columns
// colOff[r] .. colOff[r+1]-1), so clusters never mix replicas.
__global__ void k_b_col_assign(const MdgBox *__restrict__ boxes, const int *__restrict__ arep,
                               const int *__restrict__ colOff, const int *__restrict__ nfx,
                               const int *__restrict__ nfy, int n, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z,
                               int *__restrict__ colOf, int *__restrict__ colCnt,
                               float *__restrict__ zkey)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = arep[i];
    const MdgBox &B = boxes[r];
    double a = x[i] * B.u1 + y[i] * B.u2 + z[i] * B.u3 + 0.5;
    double b = x[i] * B.v1 + y[i] * B.v2 + z[i] * B.v3 + 0.5;
    double c = x[i] * B.w1 + y[i] * B.w2 + z[i] * B.w3 + 0.5;
    a -= floor(a);
    b -= floor(b);
    c -= floor(c);
    const int ncx = nfx[r], ncy = nfy[r];
    int cx = min((int)(a * ncx), ncx - 1), cy = min((int)(b * ncy), ncy - 1);
    int col = colOff[r] + cx + ncx * cy;
    colOf[i] = col;
    zkey[i] = (float)c;
    atomicAdd(&colCnt[col], 1);
}

// First cluster of each replica (and the end, at R) from the column slots.
__global__ void k_b_clu_ranges(int R, const int *__restrict__ colOff, const int *__restrict__ colSlot,
                               int *__restrict__ cluStart)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= R) cluStart[r] = colSlot[colOff[r]] / MDG_CL;
}

// Bounding box and replica of each cluster (the replica's box wraps it).
__global__ void k_b_cluster_bbox(const MdgBox *__restrict__ boxes, const int *__restrict__ arep,
                                 int nclu, const int *__restrict__ slotAtom,
                                 const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, float4 *__restrict__ bb,
                                 float4 *__restrict__ bbh, int *__restrict__ crep)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nclu) return;
    const int a0 = slotAtom[I * MDG_CL];   // a used cluster starts with an atom
    if (a0 < 0) {
        crep[I] = -1;
        bb[I] = make_float4(0.f, 0.f, 0.f, 0.f);
        bbh[I] = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
        return;
    }
    const int r = arep[a0];
    const MdgBox &B = boxes[r];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < MDG_CL; k++) {
        int a = slotAtom[I * MDG_CL + k];
        if (a < 0) continue;
        double s[3] = {x[a] * B.u1 + y[a] * B.u2 + z[a] * B.u3, x[a] * B.v1 + y[a] * B.v2 + z[a] * B.v3,
                       x[a] * B.w1 + y[a] * B.w2 + z[a] * B.w3};
        for (int d = 0; d < 3; d++) s[d] -= floor(s[d] + 0.5);
        double p[3] = {s[0] * B.a1 + s[1] * B.b1 + s[2] * B.c1, s[0] * B.a2 + s[1] * B.b2 + s[2] * B.c2,
                       s[0] * B.a3 + s[1] * B.b3 + s[2] * B.c3};
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], p[d]);
            hi[d] = fmax(hi[d], p[d]);
        }
    }
    crep[I] = r;
    bb[I] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                        (float)(0.5 * (lo[2] + hi[2])), 0.f);
    bbh[I] = make_float4((float)(0.5 * (hi[0] - lo[0])) + 1e-3f, (float)(0.5 * (hi[1] - lo[1])) + 1e-3f,
                         (float)(0.5 * (hi[2] - lo[2])) + 1e-3f, 0.f);
}

// k_cluster_pairs for a batch: J runs over I's own replica only, with that
// replica's box, and the O(N^2) predicate of verlet_list (list.c:896-1040).
__global__ void __launch_bounds__(256)
k_b_cluster_pairs(MdgParams P, const MdgBox *__restrict__ boxes, const int *__restrict__ crep,
                  const int *__restrict__ cluStart, int nclu, const int *__restrict__ slotAtom,
                  const float4 *__restrict__ bb, const float4 *__restrict__ bbh,
                  const double *__restrict__ x, const double *__restrict__ y,
                  const double *__restrict__ z, const int *__restrict__ frozen,
                  const int *__restrict__ exOff, const int *__restrict__ exLst, int kmax,
                  int *__restrict__ tmpCnt, int *__restrict__ tmpJ,
                  unsigned long long *__restrict__ tmpMask, int *__restrict__ ctrl)
{
    const int lane = threadIdx.x & 63;
    const int I = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (I >= nclu) return;
    const int r = crep[I];
    if (r < 0) {
        if (lane == 0) tmpCnt[I] = 0;
        return;
    }
    const MdgBox &B = boxes[r];
    const int jend = cluStart[r + 1];
    const float4 ci = bb[I], hi = bbh[I];
    const float L0 = (float)B.a1, L1 = (float)B.b2, L2 = (float)B.c3;
    const float rl = (float)(P.cutOff + P.delr) + 0.01f;
    const int a = lane & 7, b = lane >> 3;   // i, j atom: mask bit = lane = 8 j + i
    const int iat = slotAtom[I * MDG_CL + a];
    int found = 0;
    for (int J0 = I; J0 < jend; J0 += 64) {
        int J = J0 + lane;
        bool cand = false;
        if (J < jend) {
            float4 cj = bb[J], hj = bbh[J];
            float dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
            dx -= L0 * rintf(dx / L0);
            dy -= L1 * rintf(dy / L1);
            dz -= L2 * rintf(dz / L2);
            dx = fmaxf(fabsf(dx) - hi.x - hj.x, 0.f);
            dy = fmaxf(fabsf(dy) - hi.y - hj.y, 0.f);
            dz = fmaxf(fabsf(dz) - hi.z - hj.z, 0.f);
            cand = dx * dx + dy * dy + dz * dz <= rl * rl;
        }
        unsigned long long cm = __ballot(cand);
        while (cm) {
            int src = __ffsll((long long)cm) - 1;
            cm &= cm - 1;
            int Jc = J0 + src;
            int jat = slotAtom[Jc * MDG_CL + b];
            bool pr = false;
            if (iat >= 0 && jat >= 0 && (Jc != I || a < b))
                pr = ref_pair_pred(P, B, 0, 0, 0, 0, iat, jat, x, y, z, nullptr, nullptr, nullptr,
                                   nullptr, frozen, exOff, exLst);
            unsigned long long m = __ballot(pr);
            if (m) {
                if (found < kmax && lane == 0) {
                    tmpJ[(long)I * kmax + found] = Jc;
                    tmpMask[(long)I * kmax + found] = m;
                }
                found++;
            }
        }
    }
    if (lane == 0) {
        tmpCnt[I] = min(found, kmax);
        if (found > kmax) atomicOr(&ctrl[CTRL_LIST_OVF], 1);
    }
}

// Cluster-pair list of every replica of a batch (all O(N^2) builders).
int mdg_b_build_list(mdg_ctx *ctx)
{
    MdgBatch &T = *ctx->bt;
    const int n = ctx->n, R = T.R;
    ClusterList &C = ctx->cl;
    if (!C.colCnt) {
        C.ncol = T.hColOff[R];
        int maxslot = n + (MDG_CL - 1) * C.ncol;
        int maxclu = maxslot / MDG_CL + 1;
        HIPCHK(hipMalloc(&C.colCnt, 2 * (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colStart, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colPad, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.colSlot, (C.ncol + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.sortTmp, n * sizeof(int)));
        HIPCHK(hipMalloc(&C.sortKey, n * sizeof(float)));
        HIPCHK(hipMalloc(&C.slotAtom, ((size_t)maxclu * MDG_CL + 64) * sizeof(int)));
        HIPCHK(hipMalloc(&C.atomSlot, n * sizeof(int)));
        HIPCHK(hipMalloc(&C.bb, maxclu * sizeof(float4)));
        HIPCHK(hipMalloc(&C.bbh, maxclu * sizeof(float4)));
        HIPCHK(hipMalloc(&C.cpOff, (maxclu + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&C.tmpCnt, maxclu * sizeof(int)));
        HIPCHK(hipMalloc(&C.cluCol, maxclu * sizeof(int)));
        HIPCHK(hipMalloc(&T.cluStart, (R + 1) * sizeof(int)));
        HIPCHK(hipMalloc(&T.crep, maxclu * sizeof(int)));
        // every cluster of the largest replica, padding included
        int kmax = 64;
        for (int r = 0; r < R; r++) {
            int nr = T.hAoff[r + 1] - T.hAoff[r], ncr = T.hColOff[r + 1] - T.hColOff[r];
            kmax = std::max(kmax, (nr + (MDG_CL - 1) * ncr) / MDG_CL + 1);
        }
        C.kmax = std::max(kmax, (n + maxclu - 1) / maxclu);   // keys reuse tmpMask
        HIPCHK(hipMalloc(&C.tmpJ, (size_t)maxclu * C.kmax * sizeof(int)));
        HIPCHK(hipMalloc(&C.tmpMask, (size_t)maxclu * C.kmax * sizeof(unsigned long long)));
        C.nclu = maxclu;
        C.cap = 0;
    }
    unsigned long long *keys = (unsigned long long *)C.tmpMask;
    int *colOf = C.sortTmp;
    RETCHK(mdg_fill_i32(ctx, C.slotAtom, (long)C.nclu * MDG_CL + 64, -1));
    RETCHK(mdg_fill_i32(ctx, C.colCnt, 2 * (C.ncol + 1), 0));
    hipLaunchKernelGGL(k_b_col_assign, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, T.boxes,
                       T.arep, T.colOff, T.nfx, T.nfy, n, ctx->x, ctx->y, ctx->z, colOf, C.colCnt,
                       C.sortKey);
    LAUNCHCHK();
    RETCHK(sort_columns(ctx, colOf, keys));
    hipLaunchKernelGGL(k_b_clu_ranges, dim3(nblk(R + 1, 256)), dim3(256), 0, ctx->stream, R,
                       T.colOff, C.colSlot, T.cluStart);
    LAUNCHCHK();
    const int nclu = C.nclu;
    hipLaunchKernelGGL(k_b_cluster_bbox, dim3(nblk(nclu, 128)), dim3(128), 0, ctx->stream, T.boxes,
                       T.arep, nclu, C.slotAtom, ctx->x, ctx->y, ctx->z, C.bb, C.bbh, T.crep);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_b_cluster_pairs, dim3(nblk(nclu, 4)), dim3(256), 0, ctx->stream, ctx->P,
                       T.boxes, T.crep, T.cluStart, nclu, C.slotAtom, C.bb, C.bbh, ctx->x, ctx->y,
                       ctx->z, ctx->frozen, ctx->exOff, ctx->exLst, C.kmax, C.tmpCnt, C.tmpJ,
                       C.tmpMask, ctx->ctrl);
    LAUNCHCHK();
    RETCHK(compact_cluster_list(ctx));
    if (!T.chE) HIPCHK(hipMalloc(&T.chE, (size_t)C.maxChunks * 5 * sizeof(double)));
    return MDG_OK;
}
