// mdg_util.hip -- deterministic scans, fills and reductions on the context
// stream.  No hipMemsetAsync anywhere: every clear is a fill kernel, so the
// step sequence can be captured in a hipGraph and replayed.
#include "mdg_internal.h"

#define SCAN_BS 1024

// Block-local exclusive scan of one element per thread; returns the block
// total through *tot.
template <typename T>
__device__ inline T block_exclusive_scan(T v, T *tot)
{
    __shared__ T wsum[SCAN_BS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        T t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        const int nw = blockDim.x >> 6;
        T s = lane < nw ? wsum[lane] : T(0);
        T si = s;
        for (int o = 1; o < 64; o <<= 1) {
            T t = __shfl_up(si, o, 64);
            if (lane >= o) si += t;
        }
        if (lane < nw) wsum[lane] = si - s;   // exclusive wave offsets
        if (lane == nw - 1) *tot = si;
    }
    __syncthreads();
    return incl - v + wsum[wid];
}

template <typename Tin, typename Tout>
__global__ void k_scan_blocks(const Tin *__restrict__ in, Tout *__restrict__ out,
                              Tout *__restrict__ bsum, int n)
{
    __shared__ Tout tot;
    int i = blockIdx.x * SCAN_BS + threadIdx.x;
    Tout v = i < n ? (Tout)in[i] : Tout(0);
    Tout e = block_exclusive_scan<Tout>(v, &tot);
    if (i < n) out[i] = e;
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename Tout>
__global__ void k_scan_add(Tout *__restrict__ out, const Tout *__restrict__ boff, int n,
                           Tout *__restrict__ total_slot, const Tout *__restrict__ bsum, int nb)
{
    int i = blockIdx.x * SCAN_BS + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) *total_slot = boff[nb - 1] + bsum[nb - 1];
}

template <typename T>
__global__ void k_fill(T *p, long n, T v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Exclusive scan of up to SCAN_BS * SCAN_ONE_K elements in one block:
// thread t sums its run of k consecutive elements, the block scans the
// run sums, and each thread writes its run's exclusive prefixes; the
// total goes to out[n].  One launch instead of three to five.
#define SCAN_ONE_K 64
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(SCAN_BS) k_scan_one(const Tin *__restrict__ in,
                                                      Tout *__restrict__ out, int n, int k)
{
    __shared__ Tout tot;
    const int b = threadIdx.x * k, e = min(b + k, n);
    Tout s = Tout(0);
    for (int i = b; i < e; i++) s += (Tout)in[i];
    Tout run = block_exclusive_scan<Tout>(s, &tot);
    for (int i = b; i < e; i++) {
        const Tout v = (Tout)in[i];
        out[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) out[n] = tot;
}

// Exclusive scan with the total stored at out[n].  Scratch comes from a
// per-context pool grown on demand (never while a graph is being captured:
// the pool is sized by the first, uncaptured, call of each size).

template <typename Tin, typename Tout>
static int scan_impl(mdg_ctx *ctx, const Tin *in, Tout *out, int n, int level)
{
    if (level == 0 && n <= SCAN_BS * SCAN_ONE_K) {
        hipLaunchKernelGGL((k_scan_one<Tin, Tout>), dim3(1), dim3(SCAN_BS), 0, ctx->stream, in,
                           out, n, std::max(1, (n + SCAN_BS - 1) / SCAN_BS));
        LAUNCHCHK();
        return MDG_OK;
    }
    int nb = nblk(n, SCAN_BS);
    size_t need = (size_t)(2 * nb + 2) * sizeof(Tout);
    if (ctx->scanSize[level] < need) {
        hipStreamCaptureStatus st;
        HIPCHK(hipStreamIsCapturing(ctx->stream, &st));
        if (st != hipStreamCaptureStatusNone) {
            ctx->err = "scan scratch must be sized before graph capture";
            return MDG_EHIP;
        }
        if (ctx->scanBuf[level]) HIPCHK(hipFree(ctx->scanBuf[level]));
        HIPCHK(hipMalloc(&ctx->scanBuf[level], need * 2));
        ctx->scanSize[level] = need * 2;
    }
    Tout *bsum = (Tout *)ctx->scanBuf[level];
    Tout *boff = bsum + nb + 1;
    hipLaunchKernelGGL((k_scan_blocks<Tin, Tout>), dim3(nb), dim3(SCAN_BS), 0, ctx->stream, in,
                       out, bsum, n);
    LAUNCHCHK();
    if (nb == 1) {
        hipLaunchKernelGGL((k_fill<Tout>), dim3(1), dim3(64), 0, ctx->stream, boff, 1L, Tout(0));
        hipLaunchKernelGGL((k_scan_add<Tout>), dim3(1), dim3(SCAN_BS), 0, ctx->stream, out,
                           boff, 0, out + n, bsum, 1);
    } else {
        if (level + 1 >= 8) { ctx->err = "scan too deep"; return MDG_EARG; }
        RETCHK((scan_impl<Tout, Tout>(ctx, bsum, boff, nb, level + 1)));
        hipLaunchKernelGGL((k_scan_add<Tout>), dim3(nb), dim3(SCAN_BS), 0, ctx->stream, out,
                           boff, n, out + n, bsum, nb);
    }
    LAUNCHCHK();
    return MDG_OK;
}

template <typename T>
static int fill_impl(mdg_ctx *ctx, T *p, long n, T v)
{
    if (n <= 0) return MDG_OK;
    int nb = (int)std::min<long>(nblk(n, 256), 4096);
    hipLaunchKernelGGL((k_fill<T>), dim3(nb), dim3(256), 0, ctx->stream, p, n, v);
    LAUNCHCHK();
    return MDG_OK;
}

int mdg_fill_i64(mdg_ctx *ctx, long long *p, long n) { return fill_impl<long long>(ctx, p, n, 0); }
int mdg_fill_f64(mdg_ctx *ctx, double *p, long n, double v) { return fill_impl<double>(ctx, p, n, v); }
int mdg_fill_i32(mdg_ctx *ctx, int *p, long n, int v) { return fill_impl<int>(ctx, p, n, v); }

int mdg_scan_int_to_long(mdg_ctx *ctx, const int *in, long *out, int n)
{
    return scan_impl<int, long>(ctx, in, out, n, 0);
}

int mdg_scan_int(mdg_ctx *ctx, const int *in, int *out, int n)
{
    return scan_impl<int, int>(ctx, in, out, n, 0);
}

// Sum nblocks rows of nv partials in a fixed order (deterministic).
__global__ void k_reduce_partials(const double *__restrict__ part, int nblocks, int nv,
                                  double *__restrict__ out)
{
    int k = blockIdx.x;
    double s = 0.;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += part[(long)b * nv + k];
    __shared__ double red[16];
    s = mdg_wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        out[k] = t;
    }
}

int mdg_reduce_partials_on(hipStream_t stream, const double *part, int nblocks, int nv,
                           double *out)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(256), 0, stream, part, nblocks, nv, out);
    return hipGetLastError() == hipSuccess ? MDG_OK : MDG_EHIP;
}

int mdg_reduce_partials(mdg_ctx *ctx, const double *part, int nblocks, int nv, double *out)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(256), 0, ctx->stream, part, nblocks,
                       nv, out);
    LAUNCHCHK();
    return MDG_OK;
}
