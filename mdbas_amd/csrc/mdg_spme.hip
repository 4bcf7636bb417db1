// mdg_spme.hip -- smooth particle-mesh Ewald (spme.c:357-699) on the GPU.
//
// Per step: scaled coordinates and order-p cardinal B-splines per atom,
// charge spreading onto the m1 x m2 x m3 mesh (row-major, index 3
// fastest, as spme.c:449-491), a real-to-complex 3-D FFT (rocFFT), the
// influence function theta(k) = |b1 b2 b3|^2 exp(-|k|^2/4a^2)/|k|^2
// applied on the half spectrum with the energy and stress sums taken
// there in FP64, a complex-to-real FFT back, and the force gather.
// The reference runs full complex FFTs with the opposite exponent sign
// first (FFTW_BACKWARD then FORWARD); for a real mesh and an even theta
// the convolution is the same real field.
//
// FP64 mode spreads with FP64 atomics.  MIXED mode evaluates splines in
// FP32, spreads q*M1*M2*M3 as int64 fixed point (2^40 steps, so the mesh
// is bitwise reproducible), runs the FFTs in FP32 and gathers in FP32.
#include "mdg_internal.h"

#include <algorithm>
#include <cfloat>
#include <complex>

#define TWOPI 6.283185307179586     // global.h:53
#define PI_ 3.141592653589793
#define SQRTPI_ 1.772453850905516
#define MAXP 12

// spme.c:269-355 for one coordinate in the reference's operation order.
template <typename T>
__device__ inline void bspline_dev(double s, int p, T *bsp, T *bsd)
{
#pragma clang fp contract(off)
    const int is = (int)s;
    const T u = (T)(s - (double)is);
    bsd[0] = (T)1;
    bsd[1] = (T)-1;
    bsp[0] = u;
    bsp[1] = (T)1 - u;
    for (int k = 2; k < p; k++) {
        bsp[k] = (T)0;
        for (int j = k; j > 0; j--) {
            if (k == p - 1) bsd[j] = bsp[j] - bsp[j - 1];
            T ts;
            if (sizeof(T) == 8) ts = (T)(s + (double)j - (double)is);
            else ts = u + (T)j;
            bsp[j] = (ts * bsp[j] + ((T)(k + 1) - ts) * bsp[j - 1]) / (T)k;
        }
        if (k == p - 1) bsd[0] = bsp[0];
        bsp[0] = u * bsp[0] / (T)k;
    }
}

// bspline_dev for a compile-time order: the same operations in the same
// order, with both loops unrolled so the arrays stay in registers.
template <typename T, int P>
__device__ inline void bspline_fix(double s, T (&bsp)[P], T (&bsd)[P])
{
#pragma clang fp contract(off)
    const int is = (int)s;
    const T u = (T)(s - (double)is);
    bsd[0] = (T)1;
    bsd[1] = (T)-1;
    bsp[0] = u;
    bsp[1] = (T)1 - u;
#pragma unroll
    for (int k = 2; k < P; k++) {
        bsp[k] = (T)0;
#pragma unroll
        for (int j = k; j > 0; j--) {
            if (k == P - 1) bsd[j] = bsp[j] - bsp[j - 1];
            T ts;
            if (sizeof(T) == 8) ts = (T)(s + (double)j - (double)is);
            else ts = u + (T)j;
            bsp[j] = (ts * bsp[j] + ((T)(k + 1) - ts) * bsp[j - 1]) / (T)k;
        }
        if (k == P - 1) bsd[0] = bsp[0];
        bsp[0] = u * bsp[0] / (T)k;
    }
}

__device__ inline void scaled_coords(const MdgBox &B, int m1, int m2, int m3, double x, double y,
                                     double z, double &s1, double &s2, double &s3)
{
#pragma clang fp contract(off)
    // spme.c:434-439
    s1 = (double)m1 * (x * B.u1 + y * B.u2 + z * B.u3 + 0.5);
    s2 = (double)m2 * (x * B.v1 + y * B.v2 + z * B.v3 + 0.5);
    s3 = (double)m3 * (x * B.w1 + y * B.w2 + z * B.w3 + 0.5);
}

__device__ inline int wrap_index(int ii, int m)
{
    // spme.c:453-459
    if (ii >= m) ii = 0;
    if (ii < 0) ii += m;
    // the reference rule covers atoms within one box length of the cell;
    // anything farther is folded in rather than read out of bounds
    if (ii < 0 || ii >= m) ii = ((ii % m) + m) % m;
    return ii;
}

// Thread t spreads atom (t * stride) mod n, stride coprime to n: the atoms
// of one wavefront lie far apart, so their atomics do not collide on the
// same mesh points (neighbouring atom indices are neighbouring atoms).
// A stride near n / golden ratio with gcd(stride, n) = 1 (t -> t * stride
// mod n is then a permutation of the atoms).
static int mdg_scatter_stride(int n)
{
    if (n <= 2) return 1;
    long s = std::max(1L, (long)(n * 0.6180339887498949));
    auto gcd = [](long a, long b) {
        while (b) {
            long t = a % b;
            a = b;
            b = t;
        }
        return a;
    };
    while (gcd(s, n) != 1) s++;
    return (int)s;
}

template <bool MIXED>
__global__ void __launch_bounds__(256)
k_spme_spread(MdgParams P, MdgBox B, int n, int stride, const double *__restrict__ x,
              const double *__restrict__ y, const double *__restrict__ z,
              const double *__restrict__ q, double *grid64, long long *gridq)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int l = (int)(((long)t * stride) % n);
    const int p = P.nbsp, m1 = P.m1, m2 = P.m2, m3 = P.m3;
    double s1, s2, s3;
    scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
    typedef typename std::conditional<MIXED, float, double>::type T;
    T b1[MAXP], b2[MAXP], b3[MAXP], d[MAXP];
    bspline_dev<T>(s1, p, b1, d);
    bspline_dev<T>(s2, p, b2, d);
    bspline_dev<T>(s3, p, b3, d);
    const T ql = (T)q[l];
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    for (int i = 0; i < p; i++) {
        const int ii = wrap_index(i1 - i, m1);
        const int o1 = m2 * ii;
        const T qa = ql * b1[i];
        for (int j = 0; j < p; j++) {
            const int jj = wrap_index(i2 - j, m2);
            const long o2 = (long)m3 * (jj + o1);
            const T qb = qa * b2[j];
            for (int k = 0; k < p; k++) {
                const int kk = wrap_index(i3 - k, m3);
                if (MIXED) {
                    long long v = __double2ll_rn((double)(qb * b3[k]) * MDG_QSCALE);
                    atomicAdd((unsigned long long *)&gridq[kk + o2], (unsigned long long)v);
                } else {
                    atomicAdd(&grid64[kk + o2], (double)qb * (double)b3[k]);
                }
            }
        }
    }
}

// Row-per-atom spreading for a compile-time order P: the 16 lanes of a DPP
// row share one atom (every lane evaluates the atom's splines, a few dozen
// flops), and lane r adds the mesh rows (i, j) = (r / P, r % P), r + 16, ...
// along index 3, P points each.  The atom order is the scattered one of
// k_spme_spread.  Products are formed in the same order as k_spme_spread,
// so the mesh is the same bit for bit.
#define SPME_LPA 16   // lanes per atom

template <typename T, int P>
__device__ inline T spme_pick(const T (&v)[P], int i)
{
    T r = v[0];
#pragma unroll
    for (int k = 1; k < P; k++) r = i == k ? v[k] : r;
    return r;
}

template <bool MIXED, int P>
__global__ void __launch_bounds__(256)
k_spme_spread_rows(MdgParams Pm, MdgBox B, int n, int stride, const double *__restrict__ x,
                   const double *__restrict__ y, const double *__restrict__ z,
                   const double *__restrict__ q, double *grid64, long long *gridq)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int at = (int)(t / SPME_LPA), r = (int)(t % SPME_LPA);
    if (at >= n) return;
    const int l = (int)(((long)at * stride) % n);
    const int m1 = Pm.m1, m2 = Pm.m2, m3 = Pm.m3;
    double s1, s2, s3;
    scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
    typedef typename std::conditional<MIXED, float, double>::type T;
    T b1[P], b2[P], b3[P], d[P];
    bspline_fix<T, P>(s1, b1, d);
    bspline_fix<T, P>(s2, b2, d);
    bspline_fix<T, P>(s3, b3, d);
    const T ql = (T)q[l];
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    int kk[P];
#pragma unroll
    for (int k = 0; k < P; k++) kk[k] = wrap_index(i3 - k, m3);
#pragma unroll
    for (int ij0 = 0; ij0 < P * P; ij0 += SPME_LPA) {
        const int ij = ij0 + r;
        if (ij >= P * P) break;
        const int i = ij / P, j = ij % P;
        const long o2 = (long)m3 * (wrap_index(i2 - j, m2) + (long)m2 * wrap_index(i1 - i, m1));
        const T qb = ql * spme_pick<T, P>(b1, i) * spme_pick<T, P>(b2, j);
#pragma unroll
        for (int k = 0; k < P; k++) {
            if (MIXED) {
                long long v = __double2ll_rn((double)(qb * b3[k]) * MDG_QSCALE);
                atomicAdd((unsigned long long *)&gridq[kk[k] + o2], (unsigned long long)v);
            } else {
                atomicAdd(&grid64[kk[k] + o2], (double)qb * (double)b3[k]);
            }
        }
    }
}

// fixed-point mesh -> FP32 mesh, clearing the fixed-point mesh for the next
// step (no separate memset).
__global__ void k_grid_fix2f(long G, long long *__restrict__ gq, float *__restrict__ g32)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < G; i += (long)gridDim.x * blockDim.x) {
        g32[i] = (float)((double)gq[i] * (1.0 / MDG_QSCALE));
        gq[i] = 0;
    }
}

// Influence function on the half spectrum, energy and stress partials.
// Per block: [sum theta|F|^2, st0..st5] with full-spectrum weights.
template <typename T>
__global__ void __launch_bounds__(256)
k_spme_influence(MdgBox B, int m1, int m2, int m3, int mh, double rAlpha2,
                 const T *__restrict__ theta, T *cg, double *__restrict__ part)
{
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const long G = (long)m1 * m2 * mh;
    // grid-stride: at most mdg_infl_blocks blocks, so the record sums few rows
    for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < G;
         g += (long)gridDim.x * blockDim.x) {
        // (i, j, k) in 32-bit arithmetic (a half spectrum has < 2^31 points;
        // 64-bit division costs several times more)
        const unsigned gu = (unsigned)g, row = gu / (unsigned)mh;
        const int k = (int)(gu - row * (unsigned)mh);
        const unsigned iu = row / (unsigned)m2;
        const int i = (int)iu, j = (int)(row - iu * (unsigned)m2);
        const T th = theta[g];
        const T re = cg[2 * g], im = cg[2 * g + 1];
        if (th != (T)0) {
            const double w = (k == 0 || (2 * k == m3)) ? 1.0 : 2.0;
            const double e = (double)th * ((double)re * (double)re + (double)im * (double)im);
            // spme.c:509-547 reciprocal vector for (i, j, k)
            const int ii = i > m1 / 2 ? i - m1 : i;
            const int jj = j > m2 / 2 ? j - m2 : j;
            const int kk = k > m3 / 2 ? k - m3 : k;
            const double t1 = TWOPI * ii, t2 = TWOPI * jj, t3 = TWOPI * kk;
            const double rx = t1 * B.u1 + t2 * B.v1 + t3 * B.w1;
            const double ry = t1 * B.u2 + t2 * B.v2 + t3 * B.w2;
            const double rz = t1 * B.u3 + t2 * B.v3 + t3 * B.w3;
            const double rm = rx * rx + ry * ry + rz * rz;
            const double vam = 2.0 * (1.0 / rm - rAlpha2) * e;
            acc[0] += w * e;
            acc[1] += -w * vam * rx * rx;
            acc[2] += -w * vam * rx * ry;
            acc[3] += -w * vam * rx * rz;
            acc[4] += -w * vam * ry * ry;
            acc[5] += -w * vam * ry * rz;
            acc[6] += -w * vam * rz * rz;
        }
        cg[2 * g] = re * th;
        cg[2 * g + 1] = im * th;
    }
    mdg_block_partials<7>(acc, part);
}

// spme.c:593-646: force gather; writes the reciprocal force per atom and
// per-block sums of it (for the mean removal of spme.c:648-670).
template <bool MIXED, typename TG, typename TF>
__global__ void __launch_bounds__(256)
k_spme_gather(MdgParams P, MdgBox B, int n, const double *__restrict__ x,
              const double *__restrict__ y, const double *__restrict__ z,
              const double *__restrict__ q, const TG *__restrict__ grid, double fact1,
              TF *__restrict__ frx, TF *__restrict__ fry, TF *__restrict__ frz,
              double *__restrict__ part)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0, 0, 0};
    if (l < n) {
        const int p = P.nbsp, m1 = P.m1, m2 = P.m2, m3 = P.m3;
        double s1, s2, s3;
        scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
        typedef typename std::conditional<MIXED, float, double>::type T;
        T b1[MAXP], b2[MAXP], b3[MAXP], d1[MAXP], d2[MAXP], d3[MAXP];
        bspline_dev<T>(s1, p, b1, d1);
        bspline_dev<T>(s2, p, b2, d2);
        bspline_dev<T>(s3, p, b3, d3);
        const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
        T fbx = 0, fby = 0, fbz = 0;
        for (int i = 0; i < p; i++) {
            const int ii = wrap_index(i1 - i, m1);
            const int o1 = m2 * ii;
            for (int j = 0; j < p; j++) {
                const int jj = wrap_index(i2 - j, m2);
                const long o2 = (long)m3 * (jj + o1);
                const T c1 = d1[i] * b2[j], c2 = b1[i] * d2[j], c3 = b1[i] * b2[j];
                for (int k = 0; k < p; k++) {
                    const int kk = wrap_index(i3 - k, m3);
                    const T qt = (T)grid[kk + o2];
                    fbx += qt * c1 * b3[k];
                    fby += qt * c2 * b3[k];
                    fbz += qt * c3 * d3[k];
                }
            }
        }
        const double gx = (double)fbx * m1, gy = (double)fby * m2, gz = (double)fbz * m3;
        const double c = fact1 * q[l];
        const double fxl = c * (gx * B.u1 + gy * B.v1 + gz * B.w1);
        const double fyl = c * (gx * B.u2 + gy * B.v2 + gz * B.w2);
        const double fzl = c * (gx * B.u3 + gy * B.v3 + gz * B.w3);
        frx[l] = (TF)fxl;
        fry[l] = (TF)fyl;
        frz[l] = (TF)fzl;
        acc[0] = (double)(TF)fxl;
        acc[1] = (double)(TF)fyl;
        acc[2] = (double)(TF)fzl;
    }
    mdg_block_partials<3>(acc, part);
}

// Row-per-atom gather for a compile-time order P: lane r of the atom's
// row sums its mesh rows (i, j) as in k_spme_spread_rows, the 16 partial
// sums are added across the row, and lane 0 writes the force.  One
// partial per block as k_spme_gather.
template <bool MIXED, int P, typename TG, typename TF>
__global__ void __launch_bounds__(256)
k_spme_gather_rows(MdgParams Pm, MdgBox B, int n, const double *__restrict__ x,
                   const double *__restrict__ y, const double *__restrict__ z,
                   const double *__restrict__ q, const TG *__restrict__ grid, double fact1,
                   TF *__restrict__ frx, TF *__restrict__ fry, TF *__restrict__ frz,
                   double *__restrict__ part)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(t / SPME_LPA), r = (int)(t % SPME_LPA);
    typedef typename std::conditional<MIXED, float, double>::type T;
    T fbx = 0, fby = 0, fbz = 0;
    double s1 = 0., s2 = 0., s3 = 0.;
    const int m1 = Pm.m1, m2 = Pm.m2, m3 = Pm.m3;
    if (l < n) {
        scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
        T b1[P], b2[P], b3[P], d1[P], d2[P], d3[P];
        bspline_fix<T, P>(s1, b1, d1);
        bspline_fix<T, P>(s2, b2, d2);
        bspline_fix<T, P>(s3, b3, d3);
        const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
        int kk[P];
#pragma unroll
        for (int k = 0; k < P; k++) kk[k] = wrap_index(i3 - k, m3);
#pragma unroll
        for (int ij0 = 0; ij0 < P * P; ij0 += SPME_LPA) {
            const int ij = ij0 + r;
            if (ij >= P * P) break;
            const int i = ij / P, j = ij % P;
            const long o2 = (long)m3 * (wrap_index(i2 - j, m2) + (long)m2 * wrap_index(i1 - i, m1));
            const T bi = spme_pick<T, P>(b1, i), bj = spme_pick<T, P>(b2, j);
            const T c1 = spme_pick<T, P>(d1, i) * bj, c2 = bi * spme_pick<T, P>(d2, j);
            const T c3 = bi * bj;
#pragma unroll
            for (int k = 0; k < P; k++) {
                const T qt = (T)grid[kk[k] + o2];
                fbx += qt * c1 * b3[k];
                fby += qt * c2 * b3[k];
                fbz += qt * c3 * d3[k];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < SPME_LPA; o <<= 1) {
        fbx += __shfl_xor(fbx, o, SPME_LPA);
        fby += __shfl_xor(fby, o, SPME_LPA);
        fbz += __shfl_xor(fbz, o, SPME_LPA);
    }
    double acc[3] = {0, 0, 0};
    if (l < n && r == 0) {
        const double gx = (double)fbx * m1, gy = (double)fby * m2, gz = (double)fbz * m3;
        const double c = fact1 * q[l];
        const double fxl = c * (gx * B.u1 + gy * B.v1 + gz * B.w1);
        const double fyl = c * (gx * B.u2 + gy * B.v2 + gz * B.w2);
        const double fzl = c * (gx * B.u3 + gy * B.v3 + gz * B.w3);
        frx[l] = (TF)fxl;
        fry[l] = (TF)fyl;
        frz[l] = (TF)fzl;
        acc[0] = (double)(TF)fxl;
        acc[1] = (double)(TF)fyl;
        acc[2] = (double)(TF)fzl;
    }
    mdg_block_partials<3>(acc, part);
}

// Binned spreading without global atomics (mixed).  Every step the atoms
// go into buckets of up to SPME_BCAP atoms, one per bin of SPME_BIN^3 mesh
// points, by the top mesh index of their spline support (k_spme_bin).
// One block then owns a tile of SPME_TILE^3 mesh points: it reads the
// atoms of the bins whose support can reach the tile, adds their
// fixed-point contributions inside the tile in LDS (integer, so the sum
// does not depend on the order), and writes the tile's FP32 mesh values.
// Each contribution is formed as in k_spme_spread, so the mesh is the
// same bit for bit, and no global clearing or conversion pass is needed;
// the gather clears the bucket counts for the next step.
//
// An atom's P mesh indices per dimension are the run top, top - 1, ...,
// top - P + 1 (mod m) with top = wrap_index(i, m) whenever the spline base
// i lies in [0, m].  An atom outside that (more than one mesh spacing out
// of the box), or one that finds its bucket full, goes to a short list
// every tile checks point by point.
#define SPME_BIN 4
#define SPME_TILE 8
#define SPME_BCAP 32

__device__ inline int spme_top(int i, int m) { return (i >= 0 && i <= m) ? (i == m ? 0 : i) : -1; }

// FP32 splines of one atom for a compile-time order, as stored by
// k_spme_bin: nq float4 per dimension, values then derivatives.
template <int P>
struct SplineAtom {
    static constexpr int NQ = (P + 3) / 4;
    float b[3][4 * NQ], d[3][4 * NQ];
    int base[3];
};

template <int P>
__device__ inline void spline_load(const int4 *__restrict__ splI, const float4 *__restrict__ splB,
                                   const float4 *__restrict__ splD, int n, int l, bool deriv,
                                   SplineAtom<P> &S)
{
    constexpr int NQ = SplineAtom<P>::NQ;
    const int4 bi = splI[l];
    S.base[0] = bi.x;
    S.base[1] = bi.y;
    S.base[2] = bi.z;
#pragma unroll
    for (int dim = 0; dim < 3; dim++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const float4 v = splB[(long)(dim * NQ + q) * n + l];
            S.b[dim][4 * q] = v.x;
            S.b[dim][4 * q + 1] = v.y;
            S.b[dim][4 * q + 2] = v.z;
            S.b[dim][4 * q + 3] = v.w;
            if (deriv) {
                const float4 w = splD[(long)(dim * NQ + q) * n + l];
                S.d[dim][4 * q] = w.x;
                S.d[dim][4 * q + 1] = w.y;
                S.d[dim][4 * q + 2] = w.z;
                S.d[dim][4 * q + 3] = w.w;
            }
        }
}

// Per atom: the scaled coordinates, the FP32 splines and derivatives
// (stored once for the spread and the gather), and the bucket of the top
// mesh index (or the outlier list).
template <int P>
__global__ void __launch_bounds__(256)
k_spme_bin(MdgParams Pm, MdgBox B, int n, int nby, int nbz, const double *__restrict__ x,
           const double *__restrict__ y, const double *__restrict__ z, int *binCnt,
           int *__restrict__ bucket, int *__restrict__ binOut, int *binNOut,
           int4 *__restrict__ splI, float4 *__restrict__ splB, float4 *__restrict__ splD)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    constexpr int NQ = SplineAtom<P>::NQ;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    double s[3];
    scaled_coords(B, Pm.m1, Pm.m2, Pm.m3, x[l], y[l], z[l], s[0], s[1], s[2]);
    const int i1 = (int)s[0], i2 = (int)s[1], i3 = (int)s[2];
    // the FP64 spread evaluates its splines in FP64 itself: bins only
    if (splB) splI[l] = make_int4(i1, i2, i3, 0);
#pragma unroll
    for (int dim = 0; dim < 3 && splB; dim++) {
        float b[P], d[P], bq[4 * NQ], dq[4 * NQ];
        bspline_fix<float, P>(s[dim], b, d);
#pragma unroll
        for (int k = 0; k < 4 * NQ; k++) {
            bq[k] = k < P ? b[k] : 0.f;
            dq[k] = k < P ? d[k] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            splB[(long)(dim * NQ + q) * n + l] =
                make_float4(bq[4 * q], bq[4 * q + 1], bq[4 * q + 2], bq[4 * q + 3]);
            splD[(long)(dim * NQ + q) * n + l] =
                make_float4(dq[4 * q], dq[4 * q + 1], dq[4 * q + 2], dq[4 * q + 3]);
        }
    }
    const int t1 = spme_top(i1, Pm.m1), t2 = spme_top(i2, Pm.m2), t3 = spme_top(i3, Pm.m3);
    if (t1 >= 0 && t2 >= 0 && t3 >= 0) {
        const int b = ((t1 / SPME_BIN) * nby + t2 / SPME_BIN) * nbz + t3 / SPME_BIN;
        const int r = atomicAdd(&binCnt[b], 1);
        if (r < SPME_BCAP) {
            bucket[(long)b * SPME_BCAP + r] = l;
            return;
        }
    }
    binOut[atomicAdd(&binNOut[0], 1)] = l;
}

// The bins (at most SPME_TILE / SPME_BIN + 1 + (P - 1) / SPME_BIN + 1 per
// dimension) holding atoms whose top index lies in [X0, X0 + T + P - 2]
// (mod m); the range is shorter than m - SPME_BIN, so no bin repeats.
template <int P>
__device__ inline int spme_tile_bins(int X0, int m, int *bins)
{
    int nb = 0, last = -1;
    for (int d = 0; d < SPME_TILE + P - 1; d++) {
        const int top = (X0 + d) % m, b = top / SPME_BIN;
        if (b != last) {
            bins[nb++] = b;
            last = b;
        }
    }
    return nb;
}

// F64: the FP64 mesh of the FP64 path.  Each bucketed atom's splines are
// evaluated in FP64 from its position (as k_spme_spread_rows<false> does,
// the same products in the same order) and added into the tile with LDS
// FP64 atomics; every mesh point is written once, so the mesh needs no
// clearing.  The order of the additions varies from run to run, as it does
// with the global FP64 atomics this replaces.
template <int P, bool F64>
__global__ void __launch_bounds__(256)
k_spme_spread_tiles(MdgParams Pm, MdgBox B, int nty, int ntz, int nby, int nbz,
                    const double *__restrict__ x, const double *__restrict__ y,
                    const double *__restrict__ z, const double *__restrict__ q,
                    const int *__restrict__ binCnt, const int *__restrict__ bucket,
                    const int *__restrict__ binOut, const int *__restrict__ binNOut,
                    const int4 *__restrict__ splI, const float4 *__restrict__ splB,
                    float *__restrict__ grid32, double *__restrict__ grid64)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    constexpr int T = SPME_TILE, NBD = (SPME_TILE + P - 2) / SPME_BIN + 2;
    constexpr int NB3 = NBD * NBD * NBD;
    static_assert(NB3 <= 128, "bin list of a tile exceeds the scan width");
    __shared__ unsigned long long acc[T * T * T];
    __shared__ int bx[NBD], by[NBD], bz[NBD], nbd[3];
    __shared__ int off[NB3 + 1], bid[NB3];
    const int m1 = Pm.m1, m2 = Pm.m2, m3 = Pm.m3;
    const int tile = blockIdx.x;
    const int tx = tile / (nty * ntz), ty = (tile / ntz) % nty, tz = tile % ntz;
    const int X0 = tx * T, Y0 = ty * T, Z0 = tz * T;
    for (int i = threadIdx.x; i < T * T * T; i += blockDim.x) acc[i] = 0ull;
    if (threadIdx.x == 0) nbd[0] = spme_tile_bins<P>(X0, m1, bx);
    if (threadIdx.x == 1) nbd[1] = spme_tile_bins<P>(Y0, m2, by);
    if (threadIdx.x == 2) nbd[2] = spme_tile_bins<P>(Z0, m3, bz);
    __syncthreads();
    const int nbt = nbd[0] * nbd[1] * nbd[2];
    if (threadIdx.x < 64) {
        // the first wave: bucket sizes of the tile's bins (two per lane) and
        // their exclusive prefix sums
        const int lane = threadIdx.x;
        int cnt[2] = {0, 0};
        for (int h = 0; h < 2; h++) {
            const int c = lane + 64 * h;
            if (c < nbt) {
                const int ix = c / (nbd[1] * nbd[2]), iy = (c / nbd[2]) % nbd[1], iz = c % nbd[2];
                const int b = (bx[ix] * nby + by[iy]) * nbz + bz[iz];
                bid[c] = b;
                cnt[h] = min(binCnt[b], SPME_BCAP);
            }
        }
        int run = 0;
        for (int h = 0; h < 2; h++) {
            int v = cnt[h];
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o, 64);
                if (lane >= o) v += u;
            }
            const int c = lane + 64 * h;
            if (c < nbt) off[c] = run + v - cnt[h];
            run += __shfl(v, 63, 64);
        }
        if (lane == 0) off[nbt] = run;
    }
    __syncthreads();
    const int ncand = off[nbt], nout = binNOut[0];
    for (int c0 = 0; c0 < ncand + nout; c0 += blockDim.x) {
        const int c = c0 + threadIdx.x;
        if (c < ncand + nout) {
            int l;
            if (c < ncand) {
                int lo = 0, hi = nbt;   // off[lo] <= c < off[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (off[mid] <= c) lo = mid;
                    else hi = mid;
                }
                l = bucket[(long)bid[lo] * SPME_BCAP + (c - off[lo])];
            } else {
                l = binOut[c - ncand];
            }
            if (F64) {
#pragma clang fp contract(off)
                double s1, s2, s3;
                scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
                double b1[P], b2[P], b3[P], d[P];
                bspline_fix<double, P>(s1, b1, d);
                bspline_fix<double, P>(s2, b2, d);
                bspline_fix<double, P>(s3, b3, d);
                const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
                const double ql = q[l];
                double *acc64 = reinterpret_cast<double *>(acc);
#pragma unroll
                for (int i = 0; i < P; i++) {
                    const int lx = wrap_index(i1 - i, m1) - X0;
                    if (lx < 0 || lx >= T) continue;
                    const double qa = ql * b1[i];
#pragma unroll
                    for (int j = 0; j < P; j++) {
                        const int ly = wrap_index(i2 - j, m2) - Y0;
                        if (ly < 0 || ly >= T) continue;
                        const double qb = qa * b2[j];
#pragma unroll
                        for (int k = 0; k < P; k++) {
                            const int lz = wrap_index(i3 - k, m3) - Z0;
                            if (lz < 0 || lz >= T) continue;
                            atomicAdd(&acc64[(lx * T + ly) * T + lz], qb * b3[k]);
                        }
                    }
                }
            } else {
                SplineAtom<P> S;
                spline_load<P>(splI, splB, nullptr, Pm.n, l, false, S);
                const float ql = (float)q[l];
#pragma unroll
                for (int i = 0; i < P; i++) {
                    const int lx = wrap_index(S.base[0] - i, m1) - X0;
                    if (lx < 0 || lx >= T) continue;
                    const float qa = ql * S.b[0][i];
#pragma unroll
                    for (int j = 0; j < P; j++) {
                        const int ly = wrap_index(S.base[1] - j, m2) - Y0;
                        if (ly < 0 || ly >= T) continue;
                        const float qb = qa * S.b[1][j];
#pragma unroll
                        for (int k = 0; k < P; k++) {
                            const int lz = wrap_index(S.base[2] - k, m3) - Z0;
                            if (lz < 0 || lz >= T) continue;
                            const long long v =
                                __double2ll_rn((double)(qb * S.b[2][k]) * MDG_QSCALE);
                            atomicAdd(&acc[(lx * T + ly) * T + lz], (unsigned long long)v);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < T * T * T; i += blockDim.x) {
        const int gx = X0 + i / (T * T), gy = Y0 + (i / T) % T, gz = Z0 + i % T;
        if (gx < m1 && gy < m2 && gz < m3) {
            const long o = gz + (long)m3 * (gy + (long)m2 * gx);
            if (F64) grid64[o] = reinterpret_cast<const double *>(acc)[i];
            else grid32[o] = (float)((double)(long long)acc[i] * (1.0 / MDG_QSCALE));
        }
    }
}

// Whether the binned path covers this mesh and order (else the atomic one).
static bool spme_binned(const mdg_ctx *ctx)
{
    const int p = ctx->P.nbsp, need = SPME_TILE + p + SPME_BIN;
    return (p == 4 || p == 6 || p == 8) && ctx->m1 >= need && ctx->m2 >= need && ctx->m3 >= need;
}

static int spme_bin_alloc(mdg_ctx *ctx)
{
    if (ctx->binCnt || !spme_binned(ctx)) return MDG_OK;
    ctx->nbx = (ctx->m1 + SPME_BIN - 1) / SPME_BIN;
    ctx->nby = (ctx->m2 + SPME_BIN - 1) / SPME_BIN;
    ctx->nbz = (ctx->m3 + SPME_BIN - 1) / SPME_BIN;
    const long nb = (long)ctx->nbx * ctx->nby * ctx->nbz;
    const int n = ctx->n;
    HIPCHK(hipMalloc(&ctx->binCnt, nb * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binBucket, nb * SPME_BCAP * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binOut, n * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binNOut, 2 * sizeof(int)));
    RETCHK(mdg_fill_i32(ctx, ctx->binCnt, nb, 0));
    RETCHK(mdg_fill_i32(ctx, ctx->binNOut, 2, 0));
    if (ctx->precision == MDG_MIXED) {
        const int nq = (ctx->P.nbsp + 3) / 4;
        HIPCHK(hipMalloc(&ctx->splI, n * sizeof(int4)));
        HIPCHK(hipMalloc(&ctx->splB, 3 * nq * (size_t)n * sizeof(float4)));
        HIPCHK(hipMalloc(&ctx->splD, 3 * nq * (size_t)n * sizeof(float4)));
    }
    RETCHK(mdg_fill_i32(ctx, ctx->binCnt, nb, 0));
    RETCHK(mdg_fill_i32(ctx, ctx->binNOut, 2, 0));
    return MDG_OK;
}

// Clears the bucket counts and the outlier count for the next step.
__global__ void k_bin_clear(long nbins, int *__restrict__ binCnt, int *__restrict__ binNOut)
{
    for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < nbins;
         b += (long)gridDim.x * blockDim.x)
        binCnt[b] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) binNOut[0] = 0;
}

// Binned spreading.  Mixed: into grid32 (replaces the atomic spread and the
// fixed-point conversion), leaving the per-atom splines for the gather.
// FP64: into grid64 (replaces the mesh clear and the global FP64 atomics);
// the bucket counts are cleared after the tiles have read them (the FP64
// gather evaluates its own splines and does not touch the bins).
template <int P, bool F64>
static int spread_binned_p(mdg_ctx *ctx)
{
    const int n = ctx->n;
    const int ntx = (ctx->m1 + SPME_TILE - 1) / SPME_TILE, nty = (ctx->m2 + SPME_TILE - 1) / SPME_TILE,
              ntz = (ctx->m3 + SPME_TILE - 1) / SPME_TILE;
    hipLaunchKernelGGL(k_spme_bin<P>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P, ctx->B,
                       n, ctx->nby, ctx->nbz, ctx->x, ctx->y, ctx->z, ctx->binCnt, ctx->binBucket,
                       ctx->binOut, ctx->binNOut, F64 ? nullptr : ctx->splI,
                       F64 ? nullptr : ctx->splB, F64 ? nullptr : ctx->splD);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_spme_spread_tiles<P, F64>), dim3(ntx * nty * ntz), dim3(256), 0,
                       ctx->stream, ctx->P, ctx->B, nty, ntz, ctx->nby, ctx->nbz, ctx->x, ctx->y,
                       ctx->z, ctx->q, ctx->binCnt, ctx->binBucket, ctx->binOut, ctx->binNOut,
                       ctx->splI, ctx->splB, ctx->grid32, ctx->grid64);
    LAUNCHCHK();
    if (F64) {
        const long nb = (long)ctx->nbx * ctx->nby * ctx->nbz;
        hipLaunchKernelGGL(k_bin_clear, dim3((int)std::min<long>(nblk(nb, 256), 1024)), dim3(256),
                           0, ctx->stream, nb, ctx->binCnt, ctx->binNOut);
        LAUNCHCHK();
    }
    return MDG_OK;
}

template <bool F64>
static int spread_binned(mdg_ctx *ctx)
{
    switch (ctx->P.nbsp) {
    case 4: return spread_binned_p<4, F64>(ctx);
    case 6: return spread_binned_p<6, F64>(ctx);
    default: return spread_binned_p<8, F64>(ctx);
    }
}

// Mixed gather after a binned spread: lane r of the atom's row reads the
// stored splines instead of evaluating them (k_spme_gather_rows otherwise).
template <int P>
__global__ void __launch_bounds__(256)
k_spme_gather_pre(MdgParams Pm, MdgBox B, int n, const double *__restrict__ q,
                  const int4 *__restrict__ splI, const float4 *__restrict__ splB,
                  const float4 *__restrict__ splD, const float *__restrict__ grid, double fact1,
                  float *__restrict__ frx, float *__restrict__ fry, float *__restrict__ frz,
                  double *__restrict__ part, int nbins, int *__restrict__ binCnt,
                  int *__restrict__ binNOut)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    // the spread tiles have read the buckets: clear them for the next step
    for (long b = t; b < nbins; b += (long)gridDim.x * blockDim.x) binCnt[b] = 0;
    if (t == 0) binNOut[0] = 0;
    const int l = (int)(t / SPME_LPA), r = (int)(t % SPME_LPA);
    float fbx = 0.f, fby = 0.f, fbz = 0.f;
    const int m1 = Pm.m1, m2 = Pm.m2, m3 = Pm.m3;
    if (l < n) {
        SplineAtom<P> S;
        spline_load<P>(splI, splB, splD, n, l, true, S);
        int kk[P];
#pragma unroll
        for (int k = 0; k < P; k++) kk[k] = wrap_index(S.base[2] - k, m3);
#pragma unroll
        for (int ij0 = 0; ij0 < P * P; ij0 += SPME_LPA) {
            const int ij = ij0 + r;
            if (ij >= P * P) break;
            const int i = ij / P, j = ij % P;
            const long o2 =
                (long)m3 * (wrap_index(S.base[1] - j, m2) + (long)m2 * wrap_index(S.base[0] - i, m1));
            float bi = S.b[0][0], bj = S.b[1][0], di = S.d[0][0], dj = S.d[1][0];
#pragma unroll
            for (int k = 1; k < P; k++) {
                bi = i == k ? S.b[0][k] : bi;
                di = i == k ? S.d[0][k] : di;
                bj = j == k ? S.b[1][k] : bj;
                dj = j == k ? S.d[1][k] : dj;
            }
            const float c1 = di * bj, c2 = bi * dj, c3 = bi * bj;
#pragma unroll
            for (int k = 0; k < P; k++) {
                const float qt = grid[kk[k] + o2];
                fbx += qt * c1 * S.b[2][k];
                fby += qt * c2 * S.b[2][k];
                fbz += qt * c3 * S.d[2][k];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < SPME_LPA; o <<= 1) {
        fbx += __shfl_xor(fbx, o, SPME_LPA);
        fby += __shfl_xor(fby, o, SPME_LPA);
        fbz += __shfl_xor(fbz, o, SPME_LPA);
    }
    double acc[3] = {0, 0, 0};
    if (l < n && r == 0) {
        const double gx = (double)fbx * m1, gy = (double)fby * m2, gz = (double)fbz * m3;
        const double c = fact1 * q[l];
        const double fxl = c * (gx * B.u1 + gy * B.v1 + gz * B.w1);
        const double fyl = c * (gx * B.u2 + gy * B.v2 + gz * B.w2);
        const double fzl = c * (gx * B.u3 + gy * B.v3 + gz * B.w3);
        frx[l] = (float)fxl;
        fry[l] = (float)fyl;
        frz[l] = (float)fzl;
        acc[0] = (double)(float)fxl;
        acc[1] = (double)(float)fyl;
        acc[2] = (double)(float)fzl;
    }
    mdg_block_partials<3>(acc, part);
}

static void gather_pre(mdg_ctx *ctx, double fact1, float *frx, float *fry, float *frz,
                       double *part)
{
    const int n = ctx->n, nbins = ctx->nbx * ctx->nby * ctx->nbz;
    const dim3 g(nblk((long)n * SPME_LPA, 256)), b(256);
    switch (ctx->P.nbsp) {
    case 4:
        hipLaunchKernelGGL(k_spme_gather_pre<4>, g, b, 0, ctx->stream, ctx->P, ctx->B, n, ctx->q,
                           ctx->splI, ctx->splB, ctx->splD, ctx->grid32, fact1, frx, fry, frz, part,
                           nbins, ctx->binCnt, ctx->binNOut);
        break;
    case 6:
        hipLaunchKernelGGL(k_spme_gather_pre<6>, g, b, 0, ctx->stream, ctx->P, ctx->B, n, ctx->q,
                           ctx->splI, ctx->splB, ctx->splD, ctx->grid32, fact1, frx, fry, frz, part,
                           nbins, ctx->binCnt, ctx->binNOut);
        break;
    default:
        hipLaunchKernelGGL(k_spme_gather_pre<8>, g, b, 0, ctx->stream, ctx->P, ctx->B, n, ctx->q,
                           ctx->splI, ctx->splB, ctx->splD, ctx->grid32, fact1, frx, fry, frz, part,
                           nbins, ctx->binCnt, ctx->binNOut);
    }
}

// Orders with a row-per-atom kernel; others use the thread-per-atom ones.
static bool spme_rows(int p) { return p == 4 || p == 6 || p == 8; }

int mdg_gather_blocks(const mdg_ctx *ctx)
{
    return spme_rows(ctx->P.nbsp) ? nblk((long)ctx->n * SPME_LPA, 256) : nblk(ctx->n, 256);
}

template <bool MIXED>
static void launch_spread(mdg_ctx *ctx, double *grid64, long long *gridq)
{
    const int n = ctx->n, st = mdg_scatter_stride(n);
    const dim3 g(nblk((long)n * SPME_LPA, 256)), b(256);
    switch (ctx->P.nbsp) {
    case 4:
        hipLaunchKernelGGL((k_spme_spread_rows<MIXED, 4>), g, b, 0, ctx->stream, ctx->P, ctx->B, n,
                           st, ctx->x, ctx->y, ctx->z, ctx->q, grid64, gridq);
        break;
    case 6:
        hipLaunchKernelGGL((k_spme_spread_rows<MIXED, 6>), g, b, 0, ctx->stream, ctx->P, ctx->B, n,
                           st, ctx->x, ctx->y, ctx->z, ctx->q, grid64, gridq);
        break;
    case 8:
        hipLaunchKernelGGL((k_spme_spread_rows<MIXED, 8>), g, b, 0, ctx->stream, ctx->P, ctx->B, n,
                           st, ctx->x, ctx->y, ctx->z, ctx->q, grid64, gridq);
        break;
    default:
        hipLaunchKernelGGL(k_spme_spread<MIXED>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                           ctx->P, ctx->B, n, st, ctx->x, ctx->y, ctx->z, ctx->q, grid64, gridq);
    }
}

template <bool MIXED, typename TG, typename TF>
static void launch_gather(mdg_ctx *ctx, const TG *grid, double fact1, TF *frx, TF *fry, TF *frz,
                          double *part)
{
    const int n = ctx->n;
    const dim3 g(nblk((long)n * SPME_LPA, 256)), b(256);
    switch (ctx->P.nbsp) {
    case 4:
        hipLaunchKernelGGL((k_spme_gather_rows<MIXED, 4, TG, TF>), g, b, 0, ctx->stream, ctx->P,
                           ctx->B, n, ctx->x, ctx->y, ctx->z, ctx->q, grid, fact1, frx, fry, frz,
                           part);
        break;
    case 6:
        hipLaunchKernelGGL((k_spme_gather_rows<MIXED, 6, TG, TF>), g, b, 0, ctx->stream, ctx->P,
                           ctx->B, n, ctx->x, ctx->y, ctx->z, ctx->q, grid, fact1, frx, fry, frz,
                           part);
        break;
    case 8:
        hipLaunchKernelGGL((k_spme_gather_rows<MIXED, 8, TG, TF>), g, b, 0, ctx->stream, ctx->P,
                           ctx->B, n, ctx->x, ctx->y, ctx->z, ctx->q, grid, fact1, frx, fry, frz,
                           part);
        break;
    default:
        hipLaunchKernelGGL((k_spme_gather<MIXED, TG, TF>), dim3(nblk(n, 256)), dim3(256), 0,
                           ctx->stream, ctx->P, ctx->B, n, ctx->x, ctx->y, ctx->z, ctx->q, grid,
                           fact1, frx, fry, frz, part);
    }
}

// dE_rec/dq_l = 2 (2 pi / V) C sum_g (theta*Q)(g) M_l(g): the convolved
// FP64 mesh gathered with the spline weights (the gradient of the
// quadratic form of spme.c:674-680 with respect to one charge).
__global__ void __launch_bounds__(256)
k_spme_potential(MdgParams P, MdgBox B, int n, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ z,
                 const double *__restrict__ grid, double scale, double *__restrict__ pot)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int p = P.nbsp, m1 = P.m1, m2 = P.m2, m3 = P.m3;
    double s1, s2, s3;
    scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
    double b1[MAXP], b2[MAXP], b3[MAXP], d[MAXP];
    bspline_dev<double>(s1, p, b1, d);
    bspline_dev<double>(s2, p, b2, d);
    bspline_dev<double>(s3, p, b3, d);
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    double phi = 0.;
    for (int i = 0; i < p; i++) {
        const int ii = wrap_index(i1 - i, m1);
        const int o1 = m2 * ii;
        for (int j = 0; j < p; j++) {
            const int jj = wrap_index(i2 - j, m2);
            const long o2 = (long)m3 * (jj + o1);
            const double c = b1[i] * b2[j];
            for (int k = 0; k < p; k++) phi += grid[wrap_index(i3 - k, m3) + o2] * c * b3[k];
        }
    }
    pot[l] = scale * phi;
}

// ---------------------------------------------------------------- host side

// spme.c:170-267: |bspc(i)|^2 moduli.
static void bsp_moduli(int m, int p, std::vector<double> &mod2)
{
    std::vector<double> bsp(p + 1, 0.0);
    bsp[0] = 0.0;
    bsp[1] = 1.0;
    for (int k = 2; k < p; k++) {
        bsp[k] = 0.0;
        for (int j = k; j > 0; j--)
            bsp[j] = ((double)j * bsp[j] + (double)(k + 1 - j) * bsp[j - 1]) / ((double)k);
    }
    std::vector<std::complex<double>> epl(m);
    epl[0] = std::complex<double>(1.0, 0.0);
    for (int i = 1; i <= m / 2; i++) {
        double ang = TWOPI * (double)i / (double)m;
        epl[i] = std::complex<double>(cos(ang), sin(ang));
        epl[m - i] = std::conj(epl[i]);
    }
    mod2.assign(m, 0.0);
    for (int i = 0; i < m; i++) {
        std::complex<double> coeff(0.0, 0.0);
        for (int k = 0; k < p - 1; k++) coeff += bsp[k + 1] * epl[((long)i * k) % m];
        std::complex<double> c = epl[((long)i * (p - 1)) % m] / coeff;
        mod2[i] = std::real(c * std::conj(c));
    }
}

// spme.c:505-583 influence function on the half spectrum (i, j, k < m3/2+1),
// |b1 b2 b3|^2 exp(-k^2/4a^2)/k^2 inside the reference's reciprocal cutoff.
static void theta_host(const MdgBox &B, int m1, int m2, int m3, int p, double alpha,
                       std::vector<double> &th)
{
    const int mh = m3 / 2 + 1;
    std::vector<double> b1, b2, b3;
    bsp_moduli(m1, p, b1);
    bsp_moduli(m2, p, b2);
    bsp_moduli(m3, p, b3);
    double rAlpha2 = -0.25 / (alpha * alpha);
    double rc = std::min((double)m1 * B.u, (double)m2 * B.v);
    rc = std::min(rc, (double)m3 * B.w);
    rc = rc * 1.05 * TWOPI;
    const double rc2 = rc * rc;
    th.assign((size_t)m1 * m2 * mh, 0.0);
    for (int i = 0; i < m1; i++) {
        int ii = i > m1 / 2 ? i - m1 : i;
        double t1 = TWOPI * (double)ii;
        for (int j = 0; j < m2; j++) {
            int jj = j > m2 / 2 ? j - m2 : j;
            double t2 = TWOPI * (double)jj;
            for (int k = 0; k < mh; k++) {
                int kk = k > m3 / 2 ? k - m3 : k;
                double t3 = TWOPI * (double)kk;
                double rx = t1 * B.u1 + t2 * B.v1 + t3 * B.w1;
                double ry = t1 * B.u2 + t2 * B.v2 + t3 * B.w2;
                double rz = t1 * B.u3 + t2 * B.v3 + t3 * B.w3;
                double rm = rx * rx + ry * ry + rz * rz;
                double v = 0.0;
                if (rm > DBL_EPSILON && rm <= rc2) v = b1[i] * b2[j] * b3[k] * exp(rAlpha2 * rm) / rm;
                th[((long)i * m2 + j) * mh + k] = v;
            }
        }
    }
}

// Real-to-complex / complex-to-real 3-D plans of count meshes laid out one
// after the other (index 3 fastest), not in place.
static int make_plans_n(mdg_ctx *ctx, bool dbl, int m1, int m2, int m3, int count,
                        rocfft_plan *fwd, rocfft_plan *bwd)
{
    size_t len[3] = {(size_t)m3, (size_t)m2, (size_t)m1};
    rocfft_precision prec = dbl ? rocfft_precision_double : rocfft_precision_single;
    if (rocfft_plan_create(fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                           prec, 3, len, count, nullptr) != rocfft_status_success ||
        rocfft_plan_create(bwd, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                           prec, 3, len, count, nullptr) != rocfft_status_success) {
        ctx->err = "rocfft_plan_create failed";
        return MDG_EHIP;
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(*fwd, &w1);
    rocfft_plan_get_work_buffer_size(*bwd, &w2);
    size_t w = std::max(w1, w2);
    if (w > ctx->fftWorkSize) {
        if (ctx->fftWork) HIPCHK(hipFree(ctx->fftWork));
        HIPCHK(hipMalloc(&ctx->fftWork, w));
        ctx->fftWorkSize = w;
        rocfft_execution_info_set_work_buffer(ctx->info, ctx->fftWork, ctx->fftWorkSize);
    }
    return MDG_OK;
}

static int make_plans(mdg_ctx *ctx, bool dbl, rocfft_plan *fwd, rocfft_plan *bwd)
{
    return make_plans_n(ctx, dbl, ctx->m1, ctx->m2, ctx->m3, 1, fwd, bwd);
}

static int rocfft_init(mdg_ctx *ctx)
{
    static bool s_setup = false;
    if (!s_setup) {
        if (rocfft_setup() != rocfft_status_success) {
            ctx->err = "rocfft_setup failed";
            return MDG_EHIP;
        }
        s_setup = true;
    }
    return MDG_OK;
}

// The influence function for the context's current box (after a barostat
// step), into the buffers mdg_spme_setup made.
int mdg_spme_rebox(mdg_ctx *ctx)
{
    const long H = (long)ctx->m1 * ctx->m2 * ctx->mh;
    std::vector<double> th;
    theta_host(ctx->B, ctx->m1, ctx->m2, ctx->m3, ctx->P.nbsp, ctx->P.alpha, th);
    std::vector<float> thf(th.begin(), th.end());
    HIPCHK(hipMemcpy(ctx->theta, th.data(), H * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->theta32, thf.data(), H * sizeof(float), hipMemcpyHostToDevice));
    return MDG_OK;
}

int mdg_spme_setup(mdg_ctx *ctx)
{
    RETCHK(rocfft_init(ctx));
    const int m1 = ctx->m1, m2 = ctx->m2, m3 = ctx->m3, p = ctx->P.nbsp;
    if (p < 2 || p > MAXP) {
        ctx->err = "spline order outside 2..12";
        return MDG_EUNSUPPORTED;
    }
    ctx->mh = m3 / 2 + 1;
    const long G = (long)m1 * m2 * m3, H = (long)m1 * m2 * ctx->mh;
    if ((long)mdg_gather_blocks(ctx) * 3 > PART_R || nblk(ctx->n, 256) * 7 > PART_R) {
        ctx->err = "mesh or atom count beyond the partial-sum buffers";
        return MDG_EUNSUPPORTED;
    }
    std::vector<double> th;
    theta_host(ctx->B, m1, m2, m3, p, ctx->P.alpha, th);
    std::vector<float> thf(th.begin(), th.end());
    HIPCHK(hipMalloc(&ctx->theta, H * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->theta32, H * sizeof(float)));
    HIPCHK(hipMemcpy(ctx->theta, th.data(), H * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->theta32, thf.data(), H * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&ctx->grid64, G * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->cgrid64, 2 * H * sizeof(double)));
    if (rocfft_execution_info_create(&ctx->info) != rocfft_status_success) {
        ctx->err = "rocfft_execution_info_create failed";
        return MDG_EHIP;
    }
    rocfft_execution_info_set_stream(ctx->info, ctx->stream);
    RETCHK(make_plans(ctx, true, &ctx->fwd, &ctx->bwd));
    if (ctx->precision == MDG_MIXED) {
        HIPCHK(hipMalloc(&ctx->grid32, G * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->cgrid32, 2 * H * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->gridq, G * sizeof(long long)));
        RETCHK(mdg_fill_i64(ctx, ctx->gridq, G));
        RETCHK(make_plans(ctx, false, &ctx->fwd32, &ctx->bwd32));
    }
    RETCHK(spme_bin_alloc(ctx));
    return MDG_OK;
}

int mdg_spme_free(mdg_ctx *ctx)
{
    if (ctx->fwd) rocfft_plan_destroy(ctx->fwd);
    if (ctx->bwd) rocfft_plan_destroy(ctx->bwd);
    if (ctx->fwd32) rocfft_plan_destroy(ctx->fwd32);
    if (ctx->bwd32) rocfft_plan_destroy(ctx->bwd32);
    if (ctx->info) rocfft_execution_info_destroy(ctx->info);
    ctx->fwd = ctx->bwd = ctx->fwd32 = ctx->bwd32 = nullptr;
    ctx->info = nullptr;
    return MDG_OK;
}

static int spme_run(mdg_ctx *ctx, bool mixed, double *partInfl, double *partGather, void *frx,
                    void *fry, void *frz)
{
    const int n = ctx->n, m1 = ctx->m1, m2 = ctx->m2, m3 = ctx->m3, mh = ctx->mh;
    const long G = (long)m1 * m2 * m3, H = (long)m1 * m2 * mh;
    const double rAlpha2 = -0.25 / (ctx->P.alpha * ctx->P.alpha);
    const double fact1 = -2.0 * (TWOPI / ctx->B.vol) * ctx->P.chargeConst;
    if (mixed) {
        if (spme_binned(ctx)) {
            RETCHK(spread_binned<false>(ctx));
        } else {
            launch_spread<true>(ctx, nullptr, ctx->gridq);
            LAUNCHCHK();
            hipLaunchKernelGGL(k_grid_fix2f, dim3((int)std::min<long>(nblk(G, 256), 8192)),
                               dim3(256), 0, ctx->stream, G, ctx->gridq, ctx->grid32);
            LAUNCHCHK();
        }
        void *in[1] = {ctx->grid32}, *out[1] = {ctx->cgrid32};
        if (rocfft_execute(ctx->fwd32, in, out, ctx->info) != rocfft_status_success) {
            ctx->err = "rocfft_execute (forward) failed";
            return MDG_EHIP;
        }
        hipLaunchKernelGGL(k_spme_influence<float>, dim3(mdg_infl_blocks(H)), dim3(256), 0, ctx->stream,
                           ctx->B, m1, m2, m3, mh, rAlpha2, ctx->theta32, ctx->cgrid32, partInfl);
        LAUNCHCHK();
        void *in2[1] = {ctx->cgrid32}, *out2[1] = {ctx->grid32};
        if (rocfft_execute(ctx->bwd32, in2, out2, ctx->info) != rocfft_status_success) {
            ctx->err = "rocfft_execute (inverse) failed";
            return MDG_EHIP;
        }
        if (spme_binned(ctx))
            gather_pre(ctx, fact1, (float *)frx, (float *)fry, (float *)frz, partGather);
        else
            launch_gather<true>(ctx, (const float *)ctx->grid32, fact1, (float *)frx,
                                (float *)fry, (float *)frz, partGather);
        LAUNCHCHK();
    } else {
        if (spme_binned(ctx)) {
            RETCHK(spread_binned<true>(ctx));
        } else {
            RETCHK(mdg_fill_f64(ctx, ctx->grid64, G, 0.0));
            launch_spread<false>(ctx, ctx->grid64, nullptr);
            LAUNCHCHK();
        }
        void *in[1] = {ctx->grid64}, *out[1] = {ctx->cgrid64};
        if (rocfft_execute(ctx->fwd, in, out, ctx->info) != rocfft_status_success) {
            ctx->err = "rocfft_execute (forward) failed";
            return MDG_EHIP;
        }
        hipLaunchKernelGGL(k_spme_influence<double>, dim3(mdg_infl_blocks(H)), dim3(256), 0,
                           ctx->stream, ctx->B, m1, m2, m3, mh, rAlpha2, ctx->theta, ctx->cgrid64,
                           partInfl);
        LAUNCHCHK();
        void *in2[1] = {ctx->cgrid64}, *out2[1] = {ctx->grid64};
        if (rocfft_execute(ctx->bwd, in2, out2, ctx->info) != rocfft_status_success) {
            ctx->err = "rocfft_execute (inverse) failed";
            return MDG_EHIP;
        }
        launch_gather<false>(ctx, (const double *)ctx->grid64, fact1, (double *)frx,
                             (double *)fry, (double *)frz, partGather);
        LAUNCHCHK();
    }
    return MDG_OK;
}

// Device-resident SPME: per-block partials of sum theta|F|^2 and st0..5 at
// epart + PART_INFL, of the reciprocal forces at epart + PART_GATHER (the
// step's record and finalize kernels reduce them).  Reciprocal forces per
// atom go to frx64.. (FP64 mode) or frx.. (MIXED).
int mdg_spme_device(mdg_ctx *ctx)
{
    ctx->P.spmePrio = ctx->tune.spmePrio;
    const bool mixed = ctx->precision == MDG_MIXED;
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[2], ctx->stream));
    if (mixed)
        RETCHK(spme_run(ctx, true, ctx->epart + PART_INFL, ctx->epart + PART_GATHER, ctx->frx,
                        ctx->fry, ctx->frz));
    else
        RETCHK(spme_run(ctx, false, ctx->epart + PART_INFL, ctx->epart + PART_GATHER, ctx->frx64,
                        ctx->fry64, ctx->frz64));
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[3], ctx->stream));
    return MDG_OK;
}

// Drop-in spme_energy: FP64 always.  f (device, 3 arrays) += reciprocal
// forces, then the mean of the whole arrays is removed.  out[0..7]:
// energy, virial, stress[6].
__global__ void k_add_rec(int n, double *fx, double *fy, double *fz, const double *frx,
                          const double *fry, const double *frz, double *part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[3] = {0, 0, 0};
    if (i < n) {
        fx[i] += frx[i];
        fy[i] += fry[i];
        fz[i] += frz[i];
        acc[0] = fx[i];
        acc[1] = fy[i];
        acc[2] = fz[i];
    }
    mdg_block_partials<3>(acc, part);
}

__global__ void k_sub_mean(int n, double *fx, double *fy, double *fz, const double *sums)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fx[i] -= sums[0] / (double)n;
    fy[i] -= sums[1] / (double)n;
    fz[i] -= sums[2] / (double)n;
}

int mdg_spme64(mdg_ctx *ctx, double *fx, double *fy, double *fz, double eout[8])
{
    const int n = ctx->n;
    const long H = (long)ctx->m1 * ctx->m2 * ctx->mh;
    RETCHK(spme_run(ctx, false, ctx->epart + PART_INFL, ctx->epart + PART_GATHER, ctx->frx64,
                    ctx->fry64, ctx->frz64));
    RETCHK(mdg_reduce_partials(ctx, ctx->epart + PART_INFL, mdg_infl_blocks(H), 7,
                               ctx->scal + S_EREC));
    hipLaunchKernelGGL(k_add_rec, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, fx, fy, fz,
                       ctx->frx64, ctx->fry64, ctx->frz64, ctx->epart + PART_MISC);
    LAUNCHCHK();
    RETCHK(mdg_reduce_partials(ctx, ctx->epart + PART_MISC, nblk(n, 256), 3, ctx->scal + S_FRX));
    hipLaunchKernelGGL(k_sub_mean, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, fx, fy, fz,
                       ctx->scal + S_FRX);
    LAUNCHCHK();
    double s[7];
    HIPCHK(hipMemcpyAsync(s, ctx->scal + S_EREC, 7 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // spme.c:674-697
    const double fact = (TWOPI / ctx->B.vol) * ctx->P.chargeConst;
    const double eRec = s[0];
    double st[6];
    st[0] = fact * (s[1] + eRec) + ctx->eNet;
    st[1] = fact * s[2];
    st[2] = fact * s[3];
    st[3] = fact * (s[4] + eRec) + ctx->eNet;
    st[4] = fact * s[5];
    st[5] = fact * (s[6] + eRec) + ctx->eNet;
    eout[0] = fact * eRec + ctx->eSelf + ctx->eNet;
    eout[1] = -(st[0] + st[3] + st[5]);
    for (int k = 0; k < 6; k++) eout[2 + k] = st[k];
    return MDG_OK;
}

// dE_rec/dq per atom (FP64, self and net-charge terms not included) into
// pot (device, n); the FP64 reciprocal forces land in frx64.. as a side
// effect.
int mdg_spme_pot(mdg_ctx *ctx, double *pot)
{
    RETCHK(spme_run(ctx, false, ctx->epart + PART_INFL, ctx->epart + PART_MISC, ctx->frx64,
                    ctx->fry64, ctx->frz64));
    const double scale = 2.0 * (TWOPI / ctx->B.vol) * ctx->P.chargeConst;
    hipLaunchKernelGGL(k_spme_potential, dim3(nblk(ctx->n, 256)), dim3(256), 0, ctx->stream,
                       ctx->P, ctx->B, ctx->n, ctx->x, ctx->y, ctx->z, ctx->grid64, scale, pot);
    LAUNCHCHK();
    return MDG_OK;
}

// ---------------------------------------------------------------- second derivatives

// Spline values, first and second derivatives along one axis in bspgen's
// layout: b[i] = M_p(u+i), d[i] = M_p'(u+i), dd[i] = M_p''(u+i) =
// M_{p-2}(u+i) - 2 M_{p-2}(u+i-1) + M_{p-2}(u+i-2).
__device__ inline void hvp_splines(double s, int p, double *b, double *d, double *dd)
{
    double b2[MAXP], t[MAXP];
    bspline_dev<double>(s, p, b, d);
    bspline_dev<double>(s, p - 2, b2, t);
    for (int i = 0; i < p; i++) {
        const double v0 = i < p - 2 ? b2[i] : 0.;
        const double v1 = (i >= 1 && i - 1 < p - 2) ? b2[i - 1] : 0.;
        const double v2 = (i >= 2 && i - 2 < p - 2) ? b2[i - 2] : 0.;
        dd[i] = v0 - 2. * v1 + v2;
    }
}

__device__ inline void hvp_u(const double *u, const double *s, int n, int l, double ul[3])
{
    if (u) {
        ul[0] = u[l];
        ul[1] = u[n + l];
        ul[2] = u[2 * (long)n + l];
    } else {
        ul[0] = s[0];
        ul[1] = s[1];
        ul[2] = s[2];
    }
}

// Q_u(g) = sum_l q_l grad W_l(g) . u_l onto the FP64 mesh (atomics): the
// charge spread of spme.c:449-491 with each spline product replaced by its
// derivative along u_l.
__global__ void __launch_bounds__(256)
k_hvp_spread(MdgParams P, MdgBox B, int n, const double *__restrict__ x,
             const double *__restrict__ y, const double *__restrict__ z,
             const double *__restrict__ q, const double *__restrict__ u, HvpVec sv, double *grid)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int p = P.nbsp, m1 = P.m1, m2 = P.m2, m3 = P.m3;
    double s1, s2, s3;
    scaled_coords(B, m1, m2, m3, x[l], y[l], z[l], s1, s2, s3);
    double b1[MAXP], b2[MAXP], b3[MAXP], d1[MAXP], d2[MAXP], d3[MAXP];
    bspline_dev<double>(s1, p, b1, d1);
    bspline_dev<double>(s2, p, b2, d2);
    bspline_dev<double>(s3, p, b3, d3);
    double ul[3];
    hvp_u(u, sv.s, n, l, ul);
    const double ua = m1 * (B.u1 * ul[0] + B.u2 * ul[1] + B.u3 * ul[2]);
    const double ub = m2 * (B.v1 * ul[0] + B.v2 * ul[1] + B.v3 * ul[2]);
    const double uc = m3 * (B.w1 * ul[0] + B.w2 * ul[1] + B.w3 * ul[2]);
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    for (int i = 0; i < p; i++) {
        const int o1 = m2 * wrap_index(i1 - i, m1);
        for (int j = 0; j < p; j++) {
            const long o2 = (long)m3 * (wrap_index(i2 - j, m2) + o1);
            for (int k = 0; k < p; k++) {
                const double v = q[l] * (d1[i] * b2[j] * b3[k] * ua + b1[i] * d2[j] * b3[k] * ub +
                                         b1[i] * b2[j] * d3[k] * uc);
                atomicAdd(&grid[wrap_index(i3 - k, m3) + o2], v);
            }
        }
    }
}

// cg *= theta on the half spectrum (the influence function of spme.c:505-583).
__global__ void k_theta_mul(long H, const double *__restrict__ theta, double *cg)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= H) return;
    cg[2 * g] *= theta[g];
    cg[2 * g + 1] *= theta[g];
}

// Per atom, with phi = theta * Q (grid) and phi_u = theta * Q_u (gridU),
// f2 = 2 (2 pi / V) C:
//   hu_l += f2 q_l sum_g [phi (Hess W_l) u_l + phi_u grad W_l]
//   mq_l += f2 sum_g [phi grad W_l . u_l + phi_u W_l].
__global__ void __launch_bounds__(256)
k_hvp_gather(MdgParams P, MdgBox B, int n, const double *__restrict__ x,
             const double *__restrict__ y, const double *__restrict__ z,
             const double *__restrict__ q, const double *__restrict__ u, HvpVec sv,
             const double *__restrict__ grid, const double *__restrict__ gridU, double f2,
             double *hu, double *mq)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int p = P.nbsp, m[3] = {P.m1, P.m2, P.m3};
    double s[3];
    scaled_coords(B, m[0], m[1], m[2], x[l], y[l], z[l], s[0], s[1], s[2]);
    double b[3][MAXP], d[3][MAXP], dd[3][MAXP];
    for (int a = 0; a < 3; a++) hvp_splines(s[a], p, b[a], d[a], dd[a]);
    const double U[3][3] = {{B.u1, B.u2, B.u3}, {B.v1, B.v2, B.v3}, {B.w1, B.w2, B.w3}};
    double ul[3];
    hvp_u(u, sv.s, n, l, ul);
    double ua[3];
    for (int a = 0; a < 3; a++) ua[a] = m[a] * (U[a][0] * ul[0] + U[a][1] * ul[1] + U[a][2] * ul[2]);
    const int i0[3] = {(int)s[0], (int)s[1], (int)s[2]};
    double h[3] = {0., 0., 0.}, mql = 0.;
    for (int i = 0; i < p; i++) {
        const int o1 = m[1] * wrap_index(i0[0] - i, m[0]);
        for (int j = 0; j < p; j++) {
            const long o2 = (long)m[2] * (wrap_index(i0[1] - j, m[1]) + o1);
            for (int k = 0; k < p; k++) {
                const long g = wrap_index(i0[2] - k, m[2]) + o2;
                const double ph = grid[g], pu = gridU[g];
                const double w1[3] = {d[0][i] * b[1][j] * b[2][k], b[0][i] * d[1][j] * b[2][k],
                                      b[0][i] * b[1][j] * d[2][k]};
                double w2[3][3];
                w2[0][0] = dd[0][i] * b[1][j] * b[2][k];
                w2[1][1] = b[0][i] * dd[1][j] * b[2][k];
                w2[2][2] = b[0][i] * b[1][j] * dd[2][k];
                w2[0][1] = w2[1][0] = d[0][i] * d[1][j] * b[2][k];
                w2[0][2] = w2[2][0] = d[0][i] * b[1][j] * d[2][k];
                w2[1][2] = w2[2][1] = b[0][i] * d[1][j] * d[2][k];
                for (int a = 0; a < 3; a++) {
                    const double ca = w2[a][0] * ua[0] + w2[a][1] * ua[1] + w2[a][2] * ua[2];
                    const double c = (ph * ca + pu * w1[a]) * m[a];
                    h[0] += c * U[a][0];
                    h[1] += c * U[a][1];
                    h[2] += c * U[a][2];
                    mql += ph * w1[a] * ua[a];
                }
                mql += pu * b[0][i] * b[1][j] * b[2][k];
            }
        }
    }
    for (int c = 0; c < 3; c++) hu[(long)c * n + l] += f2 * q[l] * h[c];
    mq[l] += f2 * mql;
}

int mdg_spme_hvp(mdg_ctx *ctx, const double *u, const double s[3], double *hu, double *mq)
{
    const int n = ctx->n, m1 = ctx->m1, m2 = ctx->m2, m3 = ctx->m3, mh = ctx->mh;
    const long G = (long)m1 * m2 * m3, H = (long)m1 * m2 * mh;
    if (ctx->P.nbsp < 4) {
        ctx->err = "second derivatives of the splines need order 4 or more";
        return MDG_EUNSUPPORTED;
    }
    if (!ctx->gridU) HIPCHK(hipMalloc(&ctx->gridU, G * sizeof(double)));
    // phi = theta * Q at these positions, in grid64
    RETCHK(spme_run(ctx, false, ctx->epart + PART_INFL, ctx->epart + PART_MISC, ctx->frx64,
                    ctx->fry64, ctx->frz64));
    HvpVec sv;
    sv.s[0] = s ? s[0] : 0.;
    sv.s[1] = s ? s[1] : 0.;
    sv.s[2] = s ? s[2] : 0.;
    RETCHK(mdg_fill_f64(ctx, ctx->gridU, G, 0.0));
    hipLaunchKernelGGL(k_hvp_spread, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P, ctx->B,
                       n, ctx->x, ctx->y, ctx->z, ctx->q, u, sv, ctx->gridU);
    LAUNCHCHK();
    void *in[1] = {ctx->gridU}, *out[1] = {ctx->cgrid64};
    if (rocfft_execute(ctx->fwd, in, out, ctx->info) != rocfft_status_success) {
        ctx->err = "rocfft_execute (forward) failed";
        return MDG_EHIP;
    }
    hipLaunchKernelGGL(k_theta_mul, dim3(nblk(H, 256)), dim3(256), 0, ctx->stream, H, ctx->theta,
                       ctx->cgrid64);
    LAUNCHCHK();
    void *in2[1] = {ctx->cgrid64}, *out2[1] = {ctx->gridU};
    if (rocfft_execute(ctx->bwd, in2, out2, ctx->info) != rocfft_status_success) {
        ctx->err = "rocfft_execute (inverse) failed";
        return MDG_EHIP;
    }
    const double f2 = 2.0 * (TWOPI / ctx->B.vol) * ctx->P.chargeConst;
    hipLaunchKernelGGL(k_hvp_gather, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P, ctx->B,
                       n, ctx->x, ctx->y, ctx->z, ctx->q, u, sv, ctx->grid64, ctx->gridU, f2, hu,
                       mq);
    LAUNCHCHK();
    return MDG_OK;
}

// ---------------------------------------------------------------- replica batch

// Charge spreading for every replica at once: atom l of replica r spreads
// onto r's own cubic mesh (edge mesh[r], first element gridOff[r]) with
// r's box.
template <bool MIXED>
__global__ void __launch_bounds__(256)
k_b_spread(MdgParams P, const MdgBox *__restrict__ boxes, const int *__restrict__ arep,
           const int *__restrict__ mesh, const long *__restrict__ gridOff, int n, int stride,
           const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
           const double *__restrict__ q, double *grid64, long long *gridq)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int l = (int)(((long)t * stride) % n);   // as k_spme_spread
    const int r = arep[l];
    const MdgBox &B = boxes[r];
    const int p = P.nbsp, m = mesh[r];
    const long base = gridOff[r];
    double s1, s2, s3;
    scaled_coords(B, m, m, m, x[l], y[l], z[l], s1, s2, s3);
    typedef typename std::conditional<MIXED, float, double>::type T;
    T b1[MAXP], b2[MAXP], b3[MAXP], d[MAXP];
    bspline_dev<T>(s1, p, b1, d);
    bspline_dev<T>(s2, p, b2, d);
    bspline_dev<T>(s3, p, b3, d);
    const T ql = (T)q[l];
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    for (int i = 0; i < p; i++) {
        const int o1 = m * wrap_index(i1 - i, m);
        const T qa = ql * b1[i];
        for (int j = 0; j < p; j++) {
            const long o2 = base + (long)m * (wrap_index(i2 - j, m) + o1);
            const T qb = qa * b2[j];
            for (int k = 0; k < p; k++) {
                const int kk = wrap_index(i3 - k, m);
                if (MIXED) {
                    long long v = __double2ll_rn((double)(qb * b3[k]) * MDG_QSCALE);
                    atomicAdd((unsigned long long *)&gridq[kk + o2], (unsigned long long)v);
                } else {
                    atomicAdd(&grid64[kk + o2], (double)qb * (double)b3[k]);
                }
            }
        }
    }
}

// Influence function of replica blockIdx.x over its half spectrum, and its
// energy / stress sums (one block per replica, fixed order).  256 threads:
// 1024 halved the kernel's time alone but quadrupled it beside the pair
// kernel, which leaves few CUs with 16 free wavefront slots.
#define B_INFL_THREADS 256
template <typename T>
__global__ void __launch_bounds__(B_INFL_THREADS)
k_b_influence(const MdgBox *__restrict__ boxes, const int *__restrict__ mesh,
              const long *__restrict__ hOff, double rAlpha2, const T *__restrict__ theta, T *cg,
              double *__restrict__ repInfl)
{
    const int r = blockIdx.x;
    const MdgBox &B = boxes[r];
    const int m = mesh[r], mh = m / 2 + 1;
    const long H = (long)m * m * mh, h0 = hOff[r];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long g = threadIdx.x; g < H; g += blockDim.x) {
        // (i, j, k) in 32-bit arithmetic (a replica's half spectrum is small)
        const unsigned gu = (unsigned)g, row = gu / (unsigned)mh;
        const int k = (int)(gu - row * (unsigned)mh);
        const unsigned iu = row / (unsigned)m;
        const int i = (int)iu, j = (int)(row - iu * (unsigned)m);
        const T th = theta[h0 + g];
        T *c = cg + 2 * (h0 + g);
        const T re = c[0], im = c[1];
        if (th != (T)0) {
            const double w = (k == 0 || (2 * k == m)) ? 1.0 : 2.0;
            const double e = (double)th * ((double)re * (double)re + (double)im * (double)im);
            const int ii = i > m / 2 ? i - m : i;
            const int jj = j > m / 2 ? j - m : j;
            const int kk = k > m / 2 ? k - m : k;
            const double t1 = TWOPI * ii, t2 = TWOPI * jj, t3 = TWOPI * kk;
            const double rx = t1 * B.u1 + t2 * B.v1 + t3 * B.w1;
            const double ry = t1 * B.u2 + t2 * B.v2 + t3 * B.w2;
            const double rz = t1 * B.u3 + t2 * B.v3 + t3 * B.w3;
            const double rm = rx * rx + ry * ry + rz * rz;
            const double vam = 2.0 * (1.0 / rm - rAlpha2) * e;
            acc[0] += w * e;
            acc[1] -= w * vam * rx * rx;
            acc[2] -= w * vam * rx * ry;
            acc[3] -= w * vam * rx * rz;
            acc[4] -= w * vam * ry * ry;
            acc[5] -= w * vam * ry * rz;
            acc[6] -= w * vam * rz * rz;
        }
        c[0] = re * th;
        c[1] = im * th;
    }
    mdg_block_partials<7>(acc, repInfl);   // row blockIdx.x = r
}

// Force gather for every replica (spme.c:593-646 with r's box and mesh).
template <bool MIXED, typename TG, typename TF>
__global__ void __launch_bounds__(256)
k_b_gather(MdgParams P, const MdgBox *__restrict__ boxes, const int *__restrict__ arep,
           const int *__restrict__ mesh, const long *__restrict__ gridOff, int n,
           const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
           const double *__restrict__ q, const TG *__restrict__ grid, TF *__restrict__ frx,
           TF *__restrict__ fry, TF *__restrict__ frz)
{
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int r = arep[l];
    const MdgBox &B = boxes[r];
    const int p = P.nbsp, m = mesh[r];
    const long base = gridOff[r];
    double s1, s2, s3;
    scaled_coords(B, m, m, m, x[l], y[l], z[l], s1, s2, s3);
    typedef typename std::conditional<MIXED, float, double>::type T;
    T b1[MAXP], b2[MAXP], b3[MAXP], d1[MAXP], d2[MAXP], d3[MAXP];
    bspline_dev<T>(s1, p, b1, d1);
    bspline_dev<T>(s2, p, b2, d2);
    bspline_dev<T>(s3, p, b3, d3);
    const int i1 = (int)s1, i2 = (int)s2, i3 = (int)s3;
    T fbx = 0, fby = 0, fbz = 0;
    for (int i = 0; i < p; i++) {
        const int o1 = m * wrap_index(i1 - i, m);
        for (int j = 0; j < p; j++) {
            const long o2 = base + (long)m * (wrap_index(i2 - j, m) + o1);
            const T c1 = d1[i] * b2[j], c2 = b1[i] * d2[j], c3 = b1[i] * b2[j];
            for (int k = 0; k < p; k++) {
                const T qt = (T)grid[wrap_index(i3 - k, m) + o2];
                fbx += qt * c1 * b3[k];
                fby += qt * c2 * b3[k];
                fbz += qt * c3 * d3[k];
            }
        }
    }
    const double gx = (double)fbx * m, gy = (double)fby * m, gz = (double)fbz * m;
    const double c = -2.0 * (TWOPI / B.vol) * P.chargeConst * q[l];
    frx[l] = (TF)(c * (gx * B.u1 + gy * B.v1 + gz * B.w1));
    fry[l] = (TF)(c * (gx * B.u2 + gy * B.v2 + gz * B.w2));
    frz[l] = (TF)(c * (gx * B.u3 + gy * B.v3 + gz * B.w3));
}

// Per replica (one block each): sums of the reciprocal forces of its atoms.
template <typename TF>
__global__ void __launch_bounds__(256)
k_b_rec_sums(const int *__restrict__ aoff, const TF *__restrict__ frx, const TF *__restrict__ fry,
             const TF *__restrict__ frz, double *__restrict__ repRec)
{
    const int r = blockIdx.x;
    double acc[3] = {0, 0, 0};
    for (int i = aoff[r] + threadIdx.x; i < aoff[r + 1]; i += blockDim.x) {
        acc[0] += (double)frx[i];
        acc[1] += (double)fry[i];
        acc[2] += (double)frz[i];
    }
    mdg_block_partials<3>(acc, repRec);    // row blockIdx.x = r
}

// Binned spreading for every replica (mixed): as k_spme_bin and
// k_spme_spread_tiles, with replica r's box and cubic mesh; r's bins are
// binOff[r] .. binOff[r+1]-1 and its tiles tileOff[r] .. tileOff[r+1]-1
// (tileRep gives the replica of a tile).
template <int P>
__global__ void __launch_bounds__(256)
k_b_bin(MdgParams Pm, const MdgBox *__restrict__ boxes, const int *__restrict__ arep,
        const int *__restrict__ mesh, const int *__restrict__ binOff, int n,
        const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
        int *binCnt, int *__restrict__ bucket, int *__restrict__ binOut, int *binNOut,
        int4 *__restrict__ splI, float4 *__restrict__ splB, float4 *__restrict__ splD)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    constexpr int NQ = SplineAtom<P>::NQ;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int r = arep[l], m = mesh[r];
    double s[3];
    scaled_coords(boxes[r], m, m, m, x[l], y[l], z[l], s[0], s[1], s[2]);
    const int i1 = (int)s[0], i2 = (int)s[1], i3 = (int)s[2];
    splI[l] = make_int4(i1, i2, i3, r);
#pragma unroll
    for (int dim = 0; dim < 3; dim++) {
        float b[P], d[P], bq[4 * NQ], dq[4 * NQ];
        bspline_fix<float, P>(s[dim], b, d);
#pragma unroll
        for (int k = 0; k < 4 * NQ; k++) {
            bq[k] = k < P ? b[k] : 0.f;
            dq[k] = k < P ? d[k] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            splB[(long)(dim * NQ + q) * n + l] =
                make_float4(bq[4 * q], bq[4 * q + 1], bq[4 * q + 2], bq[4 * q + 3]);
            splD[(long)(dim * NQ + q) * n + l] =
                make_float4(dq[4 * q], dq[4 * q + 1], dq[4 * q + 2], dq[4 * q + 3]);
        }
    }
    const int t1 = spme_top(i1, m), t2 = spme_top(i2, m), t3 = spme_top(i3, m);
    if (t1 >= 0 && t2 >= 0 && t3 >= 0) {
        const int nb = (m + SPME_BIN - 1) / SPME_BIN;
        const int b = binOff[r] + ((t1 / SPME_BIN) * nb + t2 / SPME_BIN) * nb + t3 / SPME_BIN;
        const int rk = atomicAdd(&binCnt[b], 1);
        if (rk < SPME_BCAP) {
            bucket[(long)b * SPME_BCAP + rk] = l;
            return;
        }
    }
    binOut[atomicAdd(&binNOut[0], 1)] = l;
}

template <int P>
__global__ void __launch_bounds__(256)
k_b_spread_tiles(MdgParams Pm, const int *__restrict__ mesh, const int *__restrict__ binOff,
                 const int *__restrict__ tileOff, const int *__restrict__ tileRep,
                 const long *__restrict__ gridOff, const double *__restrict__ q,
                 const int *__restrict__ binCnt, const int *__restrict__ bucket,
                 const int *__restrict__ binOut, const int *__restrict__ binNOut,
                 const int4 *__restrict__ splI, const float4 *__restrict__ splB,
                 float *__restrict__ grid32)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    constexpr int T = SPME_TILE, NBD = (SPME_TILE + P - 2) / SPME_BIN + 2;
    constexpr int NB3 = NBD * NBD * NBD;
    static_assert(NB3 <= 128, "bin list of a tile exceeds the scan width");
    __shared__ unsigned long long acc[T * T * T];
    __shared__ int bx[NBD], by[NBD], bz[NBD], nbd[3];
    __shared__ int off[NB3 + 1], bid[NB3];
    const int tile = blockIdx.x, r = tileRep[tile], m = mesh[r];
    const int nt = (m + T - 1) / T, nb = (m + SPME_BIN - 1) / SPME_BIN, lt = tile - tileOff[r];
    const int tx = lt / (nt * nt), ty = (lt / nt) % nt, tz = lt % nt;
    const int X0 = tx * T, Y0 = ty * T, Z0 = tz * T;
    for (int i = threadIdx.x; i < T * T * T; i += blockDim.x) acc[i] = 0ull;
    if (threadIdx.x == 0) nbd[0] = spme_tile_bins<P>(X0, m, bx);
    if (threadIdx.x == 1) nbd[1] = spme_tile_bins<P>(Y0, m, by);
    if (threadIdx.x == 2) nbd[2] = spme_tile_bins<P>(Z0, m, bz);
    __syncthreads();
    const int nbt = nbd[0] * nbd[1] * nbd[2];
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int cnt[2] = {0, 0};
        for (int h = 0; h < 2; h++) {
            const int c = lane + 64 * h;
            if (c < nbt) {
                const int ix = c / (nbd[1] * nbd[2]), iy = (c / nbd[2]) % nbd[1], iz = c % nbd[2];
                const int b = binOff[r] + (bx[ix] * nb + by[iy]) * nb + bz[iz];
                bid[c] = b;
                cnt[h] = min(binCnt[b], SPME_BCAP);
            }
        }
        int run = 0;
        for (int h = 0; h < 2; h++) {
            int v = cnt[h];
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o, 64);
                if (lane >= o) v += u;
            }
            const int c = lane + 64 * h;
            if (c < nbt) off[c] = run + v - cnt[h];
            run += __shfl(v, 63, 64);
        }
        if (lane == 0) off[nbt] = run;
    }
    __syncthreads();
    const int ncand = off[nbt], nout = binNOut[0];
    for (int c0 = 0; c0 < ncand + nout; c0 += blockDim.x) {
        const int c = c0 + threadIdx.x;
        if (c < ncand + nout) {
            int l;
            if (c < ncand) {
                int lo = 0, hi = nbt;   // off[lo] <= c < off[hi]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (off[mid] <= c) lo = mid;
                    else hi = mid;
                }
                l = bucket[(long)bid[lo] * SPME_BCAP + (c - off[lo])];
            } else {
                l = binOut[c - ncand];
            }
            SplineAtom<P> S;
            spline_load<P>(splI, splB, nullptr, Pm.n, l, false, S);
            // the list of atoms outside the buckets is shared by all replicas
            if (c >= ncand && splI[l].w != r) continue;
            const float ql = (float)q[l];
#pragma unroll
            for (int i = 0; i < P; i++) {
                const int lx = wrap_index(S.base[0] - i, m) - X0;
                if (lx < 0 || lx >= T) continue;
                const float qa = ql * S.b[0][i];
#pragma unroll
                for (int j = 0; j < P; j++) {
                    const int ly = wrap_index(S.base[1] - j, m) - Y0;
                    if (ly < 0 || ly >= T) continue;
                    const float qb = qa * S.b[1][j];
#pragma unroll
                    for (int k = 0; k < P; k++) {
                        const int lz = wrap_index(S.base[2] - k, m) - Z0;
                        if (lz < 0 || lz >= T) continue;
                        const long long v = __double2ll_rn((double)(qb * S.b[2][k]) * MDG_QSCALE);
                        atomicAdd(&acc[(lx * T + ly) * T + lz], (unsigned long long)v);
                    }
                }
            }
        }
    }
    __syncthreads();
    const long base = gridOff[r];
    for (int i = threadIdx.x; i < T * T * T; i += blockDim.x) {
        const int gx = X0 + i / (T * T), gy = Y0 + (i / T) % T, gz = Z0 + i % T;
        if (gx < m && gy < m && gz < m)
            grid32[base + gz + (long)m * (gy + (long)m * gx)] =
                (float)((double)(long long)acc[i] * (1.0 / MDG_QSCALE));
    }
}

// Batch gather after a binned spread: the atom's row of 16 lanes reads its
// stored splines (as k_spme_gather_pre) on its replica's mesh and box.
template <int P>
__global__ void __launch_bounds__(256)
k_b_gather_pre(MdgParams Pm, const MdgBox *__restrict__ boxes, const int *__restrict__ mesh,
               const long *__restrict__ gridOff, int n, const double *__restrict__ q,
               const int4 *__restrict__ splI, const float4 *__restrict__ splB,
               const float4 *__restrict__ splD, const float *__restrict__ grid,
               float *__restrict__ frx, float *__restrict__ fry, float *__restrict__ frz)
{
    if (Pm.spmePrio == 1) __builtin_amdgcn_s_setprio(2);   // ahead of the pair kernel beside it
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = (int)(t / SPME_LPA), rr = (int)(t % SPME_LPA);
    float fbx = 0.f, fby = 0.f, fbz = 0.f;
    int r = 0, m = 1;
    if (l < n) {
        SplineAtom<P> S;
        spline_load<P>(splI, splB, splD, n, l, true, S);
        r = splI[l].w;
        m = mesh[r];
        const long base = gridOff[r];
        int kk[P];
#pragma unroll
        for (int k = 0; k < P; k++) kk[k] = wrap_index(S.base[2] - k, m);
#pragma unroll
        for (int ij0 = 0; ij0 < P * P; ij0 += SPME_LPA) {
            const int ij = ij0 + rr;
            if (ij >= P * P) break;
            const int i = ij / P, j = ij % P;
            const long o2 = base + (long)m * (wrap_index(S.base[1] - j, m) +
                                              (long)m * wrap_index(S.base[0] - i, m));
            float bi = S.b[0][0], bj = S.b[1][0], di = S.d[0][0], dj = S.d[1][0];
#pragma unroll
            for (int k = 1; k < P; k++) {
                bi = i == k ? S.b[0][k] : bi;
                di = i == k ? S.d[0][k] : di;
                bj = j == k ? S.b[1][k] : bj;
                dj = j == k ? S.d[1][k] : dj;
            }
            const float c1 = di * bj, c2 = bi * dj, c3 = bi * bj;
#pragma unroll
            for (int k = 0; k < P; k++) {
                const float qt = grid[kk[k] + o2];
                fbx += qt * c1 * S.b[2][k];
                fby += qt * c2 * S.b[2][k];
                fbz += qt * c3 * S.d[2][k];
            }
        }
    }
#pragma unroll
    for (int o = 1; o < SPME_LPA; o <<= 1) {
        fbx += __shfl_xor(fbx, o, SPME_LPA);
        fby += __shfl_xor(fby, o, SPME_LPA);
        fbz += __shfl_xor(fbz, o, SPME_LPA);
    }
    if (l < n && rr == 0) {
        const MdgBox &B = boxes[r];
        const double gx = (double)fbx * m, gy = (double)fby * m, gz = (double)fbz * m;
        const double c = -2.0 * (TWOPI / B.vol) * Pm.chargeConst * q[l];
        frx[l] = (float)(c * (gx * B.u1 + gy * B.v1 + gz * B.w1));
        fry[l] = (float)(c * (gx * B.u2 + gy * B.v2 + gz * B.w2));
        frz[l] = (float)(c * (gx * B.u3 + gy * B.v3 + gz * B.w3));
    }
}


// Whether every replica's mesh takes the binned path.
static bool b_spme_binned(const mdg_ctx *ctx)
{
    const int p = ctx->P.nbsp, need = SPME_TILE + p + SPME_BIN;
    if (!(p == 4 || p == 6 || p == 8)) return false;
    for (int m : ctx->bt->hMesh)
        if (m < need) return false;
    return true;
}

static int b_bin_alloc(mdg_ctx *ctx)
{
    MdgBatch &T = *ctx->bt;
    const int R = T.R, n = ctx->n;
    std::vector<int> bo(R + 1, 0), to(R + 1, 0), rep;
    for (int r = 0; r < R; r++) {
        const int m = T.hMesh[r], nb = (m + SPME_BIN - 1) / SPME_BIN, nt = (m + SPME_TILE - 1) / SPME_TILE;
        bo[r + 1] = bo[r] + nb * nb * nb;
        to[r + 1] = to[r] + nt * nt * nt;
        for (int k = 0; k < nt * nt * nt; k++) rep.push_back(r);
    }
    T.nTiles = to[R];
    const long nbins = bo[R];
    HIPCHK(hipMalloc(&T.binOff, (R + 1) * sizeof(int)));
    HIPCHK(hipMalloc(&T.tileOff, (R + 1) * sizeof(int)));
    HIPCHK(hipMalloc(&T.tileRep, T.nTiles * sizeof(int)));
    HIPCHK(hipMemcpy(T.binOff, bo.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T.tileOff, to.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T.tileRep, rep.data(), T.nTiles * sizeof(int), hipMemcpyHostToDevice));
    ctx->nbx = (int)nbins;   // total bins of the batch (ctx->nby = ctx->nbz = 1)
    ctx->nby = ctx->nbz = 1;
    HIPCHK(hipMalloc(&ctx->binCnt, nbins * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binBucket, nbins * SPME_BCAP * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binOut, n * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->binNOut, 2 * sizeof(int)));
    RETCHK(mdg_fill_i32(ctx, ctx->binCnt, nbins, 0));
    RETCHK(mdg_fill_i32(ctx, ctx->binNOut, 2, 0));
    const int nq = (ctx->P.nbsp + 3) / 4;
    HIPCHK(hipMalloc(&ctx->splI, n * sizeof(int4)));
    HIPCHK(hipMalloc(&ctx->splB, 3 * nq * (size_t)n * sizeof(float4)));
    HIPCHK(hipMalloc(&ctx->splD, 3 * nq * (size_t)n * sizeof(float4)));
    T.binned = true;
    return MDG_OK;
}

template <int P>
static int b_spread_binned_p(mdg_ctx *ctx)
{
    MdgBatch &T = *ctx->bt;
    const int n = ctx->n;
    hipLaunchKernelGGL(k_b_bin<P>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, ctx->P, T.boxes,
                       T.arep, T.mesh, T.binOff, n, ctx->x, ctx->y, ctx->z, ctx->binCnt,
                       ctx->binBucket, ctx->binOut, ctx->binNOut, ctx->splI, ctx->splB,
                       ctx->splD);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_b_spread_tiles<P>, dim3(T.nTiles), dim3(256), 0, ctx->stream, ctx->P,
                       T.mesh, T.binOff, T.tileOff, T.tileRep, T.gridOff, ctx->q, ctx->binCnt,
                       ctx->binBucket, ctx->binOut, ctx->binNOut, ctx->splI, ctx->splB,
                       ctx->grid32);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_bin_clear, dim3((int)std::min<long>(nblk(ctx->nbx, 256), 1024)),
                       dim3(256), 0, ctx->stream, (long)ctx->nbx, ctx->binCnt, ctx->binNOut);
    LAUNCHCHK();
    return MDG_OK;
}

static int b_gather_binned(mdg_ctx *ctx)
{
    MdgBatch &T = *ctx->bt;
    const int n = ctx->n;
    const dim3 g(nblk((long)n * SPME_LPA, 256)), b(256);
    switch (ctx->P.nbsp) {
    case 4:
        hipLaunchKernelGGL(k_b_gather_pre<4>, g, b, 0, ctx->stream, ctx->P, T.boxes, T.mesh,
                           T.gridOff, n, ctx->q, ctx->splI, ctx->splB, ctx->splD, ctx->grid32,
                           ctx->frx, ctx->fry, ctx->frz);
        break;
    case 6:
        hipLaunchKernelGGL(k_b_gather_pre<6>, g, b, 0, ctx->stream, ctx->P, T.boxes, T.mesh,
                           T.gridOff, n, ctx->q, ctx->splI, ctx->splB, ctx->splD, ctx->grid32,
                           ctx->frx, ctx->fry, ctx->frz);
        break;
    default:
        hipLaunchKernelGGL(k_b_gather_pre<8>, g, b, 0, ctx->stream, ctx->P, T.boxes, T.mesh,
                           T.gridOff, n, ctx->q, ctx->splI, ctx->splB, ctx->splD, ctx->grid32,
                           ctx->frx, ctx->fry, ctx->frz);
    }
    return MDG_OK;
}

static int b_spread_binned(mdg_ctx *ctx)
{
    switch (ctx->P.nbsp) {
    case 4: return b_spread_binned_p<4>(ctx);
    case 6: return b_spread_binned_p<6>(ctx);
    default: return b_spread_binned_p<8>(ctx);
    }
}

// Meshes, influence functions and batched FFT plans of a replica batch:
// replicas with the same mesh edge share one plan of that many transforms
// and lie next to each other in the mesh arrays.
int mdg_b_spme_setup(mdg_ctx *ctx)
{
    RETCHK(rocfft_init(ctx));
    MdgBatch &T = *ctx->bt;
    const int R = T.R, p = ctx->P.nbsp;
    if (p < 2 || p > MAXP) {
        ctx->err = "spline order outside 2..12";
        return MDG_EUNSUPPORTED;
    }
    std::vector<int> sizes(T.hMesh);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    T.hGridOff.assign(R, 0);
    T.hHOff.assign(R, 0);
    long g = 0, h = 0;
    T.fft.clear();
    for (int m : sizes) {
        MdgFftGroup F;
        F.m = m;
        F.g0 = g;
        F.h0 = h;
        for (int r = 0; r < R; r++)
            if (T.hMesh[r] == m) {
                T.hGridOff[r] = g;
                T.hHOff[r] = h;
                g += (long)m * m * m;
                h += (long)m * m * (m / 2 + 1);
                F.count++;
            }
        T.fft.push_back(F);
    }
    T.Gtot = g;
    T.Htot = h;
    std::vector<double> th(h);
    for (int r = 0; r < R; r++) {
        std::vector<double> t;
        const int m = T.hMesh[r];
        theta_host(T.hBox[r], m, m, m, p, ctx->P.alpha, t);
        std::copy(t.begin(), t.end(), th.begin() + T.hHOff[r]);
    }
    std::vector<float> thf(th.begin(), th.end());
    HIPCHK(hipMalloc(&T.gridOff, R * sizeof(long)));
    HIPCHK(hipMalloc(&T.hOff, R * sizeof(long)));
    HIPCHK(hipMemcpy(T.gridOff, T.hGridOff.data(), R * sizeof(long), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T.hOff, T.hHOff.data(), R * sizeof(long), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&ctx->theta, h * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->theta32, h * sizeof(float)));
    HIPCHK(hipMemcpy(ctx->theta, th.data(), h * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->theta32, thf.data(), h * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&ctx->grid64, g * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->cgrid64, 2 * h * sizeof(double)));
    HIPCHK(hipMalloc(&T.repInfl, R * 7 * sizeof(double)));
    HIPCHK(hipMalloc(&T.repRec, R * 3 * sizeof(double)));
    if (rocfft_execution_info_create(&ctx->info) != rocfft_status_success) {
        ctx->err = "rocfft_execution_info_create failed";
        return MDG_EHIP;
    }
    rocfft_execution_info_set_stream(ctx->info, ctx->stream);
    const bool mixed = ctx->precision == MDG_MIXED;
    for (MdgFftGroup &F : T.fft) {
        if (mixed) RETCHK(make_plans_n(ctx, false, F.m, F.m, F.m, F.count, &F.fwd32, &F.bwd32));
        else RETCHK(make_plans_n(ctx, true, F.m, F.m, F.m, F.count, &F.fwd, &F.bwd));
    }
    if (mixed) {
        HIPCHK(hipMalloc(&ctx->grid32, g * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->cgrid32, 2 * h * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->gridq, g * sizeof(long long)));
        RETCHK(mdg_fill_i64(ctx, ctx->gridq, g));
        if (b_spme_binned(ctx)) RETCHK(b_bin_alloc(ctx));
    }
    ctx->mh = 0;
    return MDG_OK;
}

void mdg_b_spme_free(mdg_ctx *ctx)
{
    for (MdgFftGroup &F : ctx->bt->fft) {
        if (F.fwd) rocfft_plan_destroy(F.fwd);
        if (F.bwd) rocfft_plan_destroy(F.bwd);
        if (F.fwd32) rocfft_plan_destroy(F.fwd32);
        if (F.bwd32) rocfft_plan_destroy(F.bwd32);
    }
    ctx->bt->fft.clear();
}

// One SPME evaluation of every replica: reciprocal forces per atom (frx..
// in MIXED, frx64.. in FP64), per-replica energy / stress sums (repInfl)
// and sums of the reciprocal forces (repRec).
int mdg_b_spme(mdg_ctx *ctx)
{
    ctx->P.spmePrio = ctx->tune.spmePrio;
    MdgBatch &T = *ctx->bt;
    const int n = ctx->n, R = T.R;
    const double rAlpha2 = -0.25 / (ctx->P.alpha * ctx->P.alpha);
    const bool mixed = ctx->precision == MDG_MIXED;
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[2], ctx->stream));
    if (mixed) {
        if (T.binned) {
            RETCHK(b_spread_binned(ctx));
        } else {
            hipLaunchKernelGGL(k_b_spread<true>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                               ctx->P, T.boxes, T.arep, T.mesh, T.gridOff, n,
                               mdg_scatter_stride(n), ctx->x, ctx->y, ctx->z, ctx->q, nullptr,
                               ctx->gridq);
            LAUNCHCHK();
            hipLaunchKernelGGL(k_grid_fix2f, dim3((int)std::min<long>(nblk(T.Gtot, 256), 8192)),
                               dim3(256), 0, ctx->stream, T.Gtot, ctx->gridq, ctx->grid32);
            LAUNCHCHK();
        }
        for (MdgFftGroup &F : T.fft) {
            void *in[1] = {ctx->grid32 + F.g0}, *out[1] = {ctx->cgrid32 + 2 * F.h0};
            if (rocfft_execute(F.fwd32, in, out, ctx->info) != rocfft_status_success) {
                ctx->err = "rocfft_execute (forward) failed";
                return MDG_EHIP;
            }
        }
        hipLaunchKernelGGL(k_b_influence<float>, dim3(R), dim3(B_INFL_THREADS), 0, ctx->stream, T.boxes,
                           T.mesh, T.hOff, rAlpha2, ctx->theta32, ctx->cgrid32, T.repInfl);
        LAUNCHCHK();
        for (MdgFftGroup &F : T.fft) {
            void *in[1] = {ctx->cgrid32 + 2 * F.h0}, *out[1] = {ctx->grid32 + F.g0};
            if (rocfft_execute(F.bwd32, in, out, ctx->info) != rocfft_status_success) {
                ctx->err = "rocfft_execute (inverse) failed";
                return MDG_EHIP;
            }
        }
        if (T.binned)
            RETCHK(b_gather_binned(ctx));
        else
            hipLaunchKernelGGL((k_b_gather<true, float, float>), dim3(nblk(n, 256)), dim3(256), 0,
                               ctx->stream, ctx->P, T.boxes, T.arep, T.mesh, T.gridOff, n,
                               ctx->x, ctx->y, ctx->z, ctx->q, ctx->grid32, ctx->frx, ctx->fry,
                               ctx->frz);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_b_rec_sums<float>, dim3(R), dim3(256), 0, ctx->stream, T.aoff,
                           ctx->frx, ctx->fry, ctx->frz, T.repRec);
        LAUNCHCHK();
    } else {
        RETCHK(mdg_fill_f64(ctx, ctx->grid64, T.Gtot, 0.0));
        hipLaunchKernelGGL(k_b_spread<false>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                           ctx->P, T.boxes, T.arep, T.mesh, T.gridOff, n, mdg_scatter_stride(n),
                           ctx->x, ctx->y, ctx->z, ctx->q, ctx->grid64, nullptr);
        LAUNCHCHK();
        for (MdgFftGroup &F : T.fft) {
            void *in[1] = {ctx->grid64 + F.g0}, *out[1] = {ctx->cgrid64 + 2 * F.h0};
            if (rocfft_execute(F.fwd, in, out, ctx->info) != rocfft_status_success) {
                ctx->err = "rocfft_execute (forward) failed";
                return MDG_EHIP;
            }
        }
        hipLaunchKernelGGL(k_b_influence<double>, dim3(R), dim3(B_INFL_THREADS), 0, ctx->stream, T.boxes,
                           T.mesh, T.hOff, rAlpha2, ctx->theta, ctx->cgrid64, T.repInfl);
        LAUNCHCHK();
        for (MdgFftGroup &F : T.fft) {
            void *in[1] = {ctx->cgrid64 + 2 * F.h0}, *out[1] = {ctx->grid64 + F.g0};
            if (rocfft_execute(F.bwd, in, out, ctx->info) != rocfft_status_success) {
                ctx->err = "rocfft_execute (inverse) failed";
                return MDG_EHIP;
            }
        }
        hipLaunchKernelGGL((k_b_gather<false, double, double>), dim3(nblk(n, 256)), dim3(256), 0,
                           ctx->stream, ctx->P, T.boxes, T.arep, T.mesh, T.gridOff, n, ctx->x,
                           ctx->y, ctx->z, ctx->q, ctx->grid64, ctx->frx64, ctx->fry64,
                           ctx->frz64);
        LAUNCHCHK();
        hipLaunchKernelGGL(k_b_rec_sums<double>, dim3(R), dim3(256), 0, ctx->stream, T.aoff,
                           ctx->frx64, ctx->fry64, ctx->frz64, T.repRec);
        LAUNCHCHK();
    }
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[3], ctx->stream));
    return MDG_OK;
}

// Runs fn (an SPME evaluation) on the context's second stream, forked from
// and joined back into the main stream by events; rocFFT follows the
// stream.  Without a second stream, or while profiling, fn runs in place.
int mdg_spme_fork(mdg_ctx *ctx)
{
    // profiled steps keep one stream, so each kernel's events time it alone
    ctx->spmeForked = ctx->stream2 && !ctx->profiling;
    if (ctx->spmeForked) HIPCHK(hipEventRecord(ctx->evFork, ctx->stream));
    return MDG_OK;
}

int mdg_spme_on(mdg_ctx *ctx, int (*fn)(mdg_ctx *))
{
    if (!ctx->spmeForked) return fn(ctx);
    hipStream_t main = ctx->stream;
    HIPCHK(hipStreamWaitEvent(ctx->stream2, ctx->evFork, 0));
    ctx->stream = ctx->stream2;
    rocfft_execution_info_set_stream(ctx->info, ctx->stream2);
    const int st = fn(ctx);
    ctx->stream = main;
    rocfft_execution_info_set_stream(ctx->info, main);
    RETCHK(st);
    HIPCHK(hipEventRecord(ctx->evJoin, ctx->stream2));
    return MDG_OK;
}
