// mdg_noalloc.hip -- mdg_nonbond_energy_noalloc: the cutoff nonbonded
// energy and forces (nonbond_energy, energy.c:216-345, the computation of
// mdg_nonbond_energy) for callers that cannot allocate device memory: code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.  Every array is the caller's device memory; the list offsets,
// the per-block partial sums and the sums themselves live in one fixed
// file-scope device array (scratch) of MDG_NOALLOC_MAX_ATOMS atoms' worth.
//
// Each GPU holds its own copy of scratch until the process ends; its
// address on the current device is looked up once per device
// (hipGetSymbolAddress) and kept.  Calls that overlap on one GPU share that
// copy and race; calls on one stream, or otherwise ordered, do not.  Every
// word a call reads is written by that call first (the offsets by the scan
// kernel, the partials by each pair block, the sums by the reduction), so
// nothing carries over from the previous call and nothing is cleared.
#include <mutex>

#include "mdg_internal.h"

namespace {
constexpr int MAX_ATOMS = MDG_NOALLOC_MAX_ATOMS;
constexpr int MAX_BLOCKS = (MAX_ATOMS + 255) / 256;
constexpr int NV = 6;   // k_pair_ref64 partials per block
// layout: list offsets (MAX_ATOMS + 1 longs), partials (MAX_BLOCKS x NV
// doubles), their sums (NV doubles): 2.1 MB
constexpr size_t OFF_BYTES = ((size_t)(MAX_ATOMS + 1) * sizeof(long) + 255) / 256 * 256;
constexpr size_t PART_BYTES = ((size_t)MAX_BLOCKS * NV * sizeof(double) + 255) / 256 * 256;
constexpr size_t SCRATCH_BYTES = OFF_BYTES + PART_BYTES + 256;
}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES];

// The address of this device's scratch (host code never uses the symbol
// itself, which is a host-side stand-in).
void *scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

// off[l + 1] = sum over m <= l of neighPair[order[m]], off[0] = 0, for
// n <= MAX_ATOMS: one block of 1024 threads, thread t scanning its own
// contiguous segment after a block scan of the segment sums.
__global__ void __launch_bounds__(1024)
k_list_offsets(int n, const int *__restrict__ neighPair, const int *__restrict__ order,
               long *__restrict__ off)
{
    __shared__ long seg[1024];
    const int t = threadIdx.x;
    const int len = (n + 1023) / 1024;
    const int l0 = min(t * len, n), l1 = min(l0 + len, n);
    long s = 0;
    for (int l = l0; l < l1; l++) s += neighPair[order[l]];
    seg[t] = s;
    __syncthreads();
    // inclusive scan of the 1024 segment sums (Hillis-Steele)
    for (int d = 1; d < 1024; d <<= 1) {
        const long v = t >= d ? seg[t - d] : 0;
        __syncthreads();
        seg[t] += v;
        __syncthreads();
    }
    long run = seg[t] - s;   // exclusive prefix of this segment
    if (t == 0) off[0] = 0;
    for (int l = l0; l < l1; l++) {
        run += neighPair[order[l]];
        off[l + 1] = run;
    }
}

// ener[0..3] += elec, vdw, virelec, virvdw (sums[0..3]), as the reference
// accumulates into ENERGY (energy.c:337-340).
__global__ void k_add_energies(const double *__restrict__ sums, double *__restrict__ ener)
{
    if (threadIdx.x < 4) ener[threadIdx.x] += sums[threadIdx.x];
}

}  // namespace

extern "C" int mdg_nonbond_energy_noalloc(const MDG_CTRL *ctrl, const MDG_PARAM *param,
                                          const MDG_PBC *box, const double *x, const double *y,
                                          const double *z, double *fx, double *fy, double *fz,
                                          const double *q, const double *eps, const double *sig,
                                          const int *neighList, const int *neighPair,
                                          const int *neighOrder, double *ener, void *stream)
{
    if (!ctrl || !param || !box || !x || !y || !z || !fx || !fy || !fz || !q || !eps || !sig ||
        !neighList || !neighPair || !neighOrder || !ener)
        return MDG_EARG;
    const int n = param->nAtom;
    if (n <= 0) return MDG_EARG;
    if (n > MAX_ATOMS) {
        fprintf(stderr,
                "mdg_nonbond_energy_noalloc: %d atoms exceed the fixed scratch of %d atoms\n", n,
                MAX_ATOMS);
        return MDG_ESCRATCH;
    }
    if (box->type < 1 || box->type > 3) return MDG_EUNSUPPORTED;
    // energy.c:45-65: NOELEC and FULL evaluate no electrostatics
    int et = ctrl->elecType;
    if (et == 0) et = 1;
    if (et < 1 || et > 4) return MDG_EUNSUPPORTED;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return MDG_EHIP;
    unsigned char *base = (unsigned char *)scratch_ptr(dev);
    if (!base) return MDG_EHIP;
    long *off = (long *)base;
    double *part = (double *)(base + OFF_BYTES);
    double *sums = (double *)(base + OFF_BYTES + PART_BYTES);
    MdgParams P;
    mdg_params_from(ctrl, param, P);
    MdgBox B;
    mdg_box_from(box, B);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_list_offsets, dim3(1), dim3(1024), 0, s, n, neighPair, neighOrder, off);
    if (hipGetLastError() != hipSuccess) return MDG_EHIP;
    const int st = mdg_pair_ref64_launch(s, P, B, et, x, y, z, q, eps, sig, neighList, off,
                                         neighOrder, fx, fy, fz, part);
    if (st) return st;
    if (mdg_reduce_partials_on(s, part, nblk(n, 256), NV, sums)) return MDG_EHIP;
    hipLaunchKernelGGL(k_add_energies, dim3(1), dim3(64), 0, s, sums, ener);
    return hipGetLastError() == hipSuccess ? MDG_OK : MDG_EHIP;
}
