// kernels.hip -- HIP kernels (gfx950) of the batched contact solve.
//
// Grid functions are stored per component (SoA), x fastest, element ii = ix + iy*mx
// (src/m_gridfunc.f90:157-181).  Element states are int8.  All kernels are batched: a launch
// serves a list of per-patch items, each owning a contiguous range of work-groups.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "fft_core.hpp"
#include "ops.hpp"
#include "kernels.hpp"

namespace cntc {

constexpr int ELEM_THREADS = 256;
constexpr int ELEM_PER_THREAD = 4;
constexpr int ROW_THREADS = 64;
constexpr int FFT_THREADS = 256;
constexpr int MAXK = 12;

// ------------------------------------------------------------------------------------------
// item lookup and reductions
// ------------------------------------------------------------------------------------------

__device__ inline int find_item(const int* gstart, int nitems, int g)
{
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (gstart[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline bool in_mask(int mask, int8_t e)
{
    switch (mask) {
    case M_ALL: return true;
    case M_INT: return e >= ADHES;
    case M_EXT: return e == EXTER;
    case M_ADH: return e == ADHES;
    case M_SLP: return e == SLIP;
    }
    return false;
}

// combine modes: 0 sum, 1 max, 2 min
__device__ inline double combine(double a, double b, int mode)
{
    return mode == 0 ? a + b : (mode == 1 ? fmax(a, b) : fmin(a, b));
}

// work-group reduction of K values; result valid in thread 0.  Wave-level steps use
// cross-lane shuffles (DPP on CDNA), then one LDS exchange between the waves.
template <int K>
__device__ inline void group_reduce(double (&v)[K], int mode, double* lds)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwave = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x = combine(x, __shfl_xor(x, off), mode);
        v[k] = x;
    }
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < K; k++) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) {
            double x = lds[k];
            for (int w = 1; w < nwave; w++) x = combine(x, lds[w * K + k], mode);
            v[k] = x;
        }
    }
}

__device__ inline double combine_identity(int mode)
{
    return mode == 0 ? 0.0 : (mode == 1 ? -INFINITY : INFINITY);
}

// The last work-group of an item to finish combines the item's group partials: one wave, lane
// l taking groups l, l + 64, ... in order, then a butterfly over the lanes -- a fixed order
// that depends only on the item's group count, so an item's result is the same whichever
// other items share the launch.
template <int K>
__device__ inline void final_reduce(const RedOut& ro, const int* gstart, int it, double* extra = nullptr)
{
    __threadfence();   // acquire: the other groups' partials
    if (threadIdx.x >= 64) return;
    const int lane = threadIdx.x, g0 = gstart[it], g1 = gstart[it + 1];
    const int mode = ro.modes[it];
    for (int k = 0; k < K; k++) {
        double x = combine_identity(mode);
        for (int g = g0 + lane; g < g1; g += 64) x = combine(x, ro.partials[(size_t)g * K + k], mode);
        for (int off = 32; off > 0; off >>= 1) x = combine(x, __shfl_xor(x, off), mode);
        if (lane == 0) {
            ro.results[(size_t)it * K + k] = x;
            if (extra) extra[k] = x;   // the item's device copy, read by later launches
        }
    }
    if (lane == 0) ro.count[it] = 0u;   // ready for the next launch that uses this counter
}

// Reduce K values over the work-group and store the group's partials; with ro.count set, the
// item's last group also forms the item's results (no separate finalize launch).  Must be
// reached by every thread of the group.
template <int K>
__device__ inline void write_partials(double (&v)[K], int mode, const RedOut& ro, const int* gstart, int it,
                                      double* extra = nullptr)
{
    __shared__ double red_lds[16 * MAXK];
    __shared__ int s_last;
    group_reduce<K>(v, mode, red_lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) ro.partials[(size_t)blockIdx.x * K + k] = v[k];
        int last = 0;
        if (ro.count) {
            __threadfence();   // release: this group's partials before its arrival
            const unsigned ng = (unsigned)(gstart[it + 1] - gstart[it]);
            last = atomicAdd(&ro.count[it], 1u) == ng - 1u;
        }
        s_last = last;
    }
    __syncthreads();
    if (s_last) final_reduce<K>(ro, gstart, it, extra);
}

// ------------------------------------------------------------------------------------------
// masked BLAS-1 (src/m_gridfunc.f90:936-1650)
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(ELEM_THREADS) k_vec(const VecItem* items, const int* gstart,
                                                      int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const VecItem& I = items[it];
    int g0 = gstart[it];
    int base = (blockIdx.x - g0) * ELEM_THREADS * ELEM_PER_THREAD;
    int mode = I.sub == V_MAXABS || I.sub == V_MAX ? 1 : (I.sub == V_MIN ? 2 : 0);
    double acc[4] = {mode == 0 ? 0.0 : (mode == 1 ? -INFINITY : INFINITY), 0.0, 0.0, 0.0};
    // the scalars formed on the device from earlier launches' results, as the host would
    double a = I.a, b = I.b;
    int sub = I.sub;
    if (I.sfn == SF_NEG_MEAN) {
        a = -I.sin[0][0] / fmax(1.0, I.sin[0][1]);
    } else if (I.sfn == SF_NCG_BETA) {
        const double rz2 = I.sin[0][0], rzr = I.sin[1][0], rz1 = I.sin[2][0];
        if (rz1 < NCG_TINY) {
            sub = V_COPY;
        } else {
            a = 1.0;
            b = fmax(0.0, (rz2 - rzr) / fmax(NCG_TINY, rz1));
        }
    } else if (I.sfn == SF_NCG_ALPHA) {
        const double rv = I.sin[0][0], vav = I.sin[1][0];
        if (fabs(vav) > 1e-32 && I.flag) a = rv / vav;
        else a = rv / fmax(NCG_TINY, vav);
    }
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (!in_mask(I.mask, I.el ? I.el[i] : (int8_t)0)) continue;
        for (int c = 0; c < I.ncomp; c++) {
            switch (sub) {
            case V_SET: I.x[c][i] = a; break;
            case V_COPY: I.y[c][i] = I.x[c][i]; break;
            case V_SCAL: I.x[c][i] *= a; break;
            case V_AXPY: I.y[c][i] += a * I.x[c][i]; break;
            case V_AXPBY: I.y[c][i] = a * I.x[c][i] + b * I.y[c][i]; break;
            case V_ADDC: I.x[c][i] += a; break;
            case V_DOT: acc[0] += I.x[c][i] * I.y[c][i]; acc[1] += 1.0; break;
            case V_SUM: acc[0] += I.x[c][i]; acc[1] += 1.0; break;
            case V_SUMSQ: acc[0] += I.x[c][i] * I.x[c][i]; acc[1] += 1.0; break;
            case V_MAXABS: acc[0] = fmax(acc[0], fabs(I.x[c][i])); break;
            case V_MIN: acc[0] = fmin(acc[0], I.x[c][i]); break;
            case V_MAX: acc[0] = fmax(acc[0], I.x[c][i]); break;
            case V_COUNT: acc[0] += 1.0; break;
            }
        }
    }
    // max/min: the second value is 0 (the count of the sums); the group reduction of
    // zeros under max or min stays 0.  The echoed scalars enter the sums once, from the
    // item's first group's first thread.
    if (mode != 0) acc[1] = 0.0;
    if (mode == 0 && I.echo[0] && blockIdx.x == gstart[it] && threadIdx.x == 0) {
        acc[2] = I.echo[0][0];
        acc[3] = I.echo[1][0];
    }
    write_partials<4>(acc, mode, ro, gstart, it, I.sout);
}

// ------------------------------------------------------------------------------------------
// NormCG and snorm element steps
// ------------------------------------------------------------------------------------------

// interior elements with negative pressure move to the exterior (src/m_solvpn.f90:302-311);
// returns (#moved, #interior after, sum pn)
__global__ void __launch_bounds__(ELEM_THREADS) k_ncg_negpn(const NcgItem* items, const int* gstart,
                                                            int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int8_t s = I.el[i];
        if (s >= ADHES && I.p[i] < 0.0) {
            I.el[i] = EXTER;
            I.p[i] = 0.0;
            s = EXTER;
            acc[0] += 1.0;
        }
        if (s >= ADHES) acc[1] += 1.0;
        acc[2] += I.p[i];
    }
    write_partials<3>(acc, 0, ro, gstart, it);
}

// elements at the deepest interpenetration enter: mode 1 at the start of NormCG with N=1
// (src/m_solvpn.f90:140-152), mode 2 when no contact is left (src/m_solvpn.f90:316-326)
__global__ void __launch_bounds__(ELEM_THREADS) k_ncg_hmin(const NcgItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        bool enter = I.mode == 1 ? I.h[i] < I.hmin : I.h[i] <= I.hmin;
        if (enter) {
            if (I.el[i] < ADHES) acc[0] += 1.0;
            I.el[i] = ADHES;
            if (I.mode == 2) I.p[i] = 0.0;
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// dd = A ps - rhs on all elements; returns (sum dd, count) over the interior
__global__ void __launch_bounds__(ELEM_THREADS) k_ncg_dd(const NcgItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double d = I.dd[i] - I.rhs[i];
        I.dd[i] = d;
        if (I.el[i] >= ADHES) { acc[0] += d; acc[1] += 1.0; }
    }
    write_partials<2>(acc, 0, ro, gstart, it);
}

// proj_avg of dd, residual, and interpenetration check of the exterior
// (src/m_solvpn.f90:351-390); returns #elements that entered
__global__ void __launch_bounds__(ELEM_THREADS) k_ncg_expand(const NcgItem* items, const int* gstart,
                                                             int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int8_t s = I.el[i];
        double d = I.dd[i];
        if (s >= ADHES) {
            d -= I.avg;
            I.dd[i] = d;
            I.res[i] = -d;
        } else {
            I.res[i] = 0.0;
            if (s == EXTER && d - I.davg < 0.0) {
                I.el[i] = ADHES;
                s = ADHES;
                I.res[i] = -(d - I.davg);
                acc[0] += 1.0;
            }
        }
        if (s == EXTER) I.v[i] = 0.0;
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// snorm contract step (src/m_snorm.f90:196-219): returns (#moved out, #small pn)
__global__ void __launch_bounds__(ELEM_THREADS) k_sn_contract(const NcgItem* items, const int* gstart,
                                                              int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] >= ADHES) {
            if (I.p[i] < -I.errpn) {
                I.el[i] = EXTER;
                I.p[i] = 0.0;
                acc[0] += 1.0;
            }
            if (fabs(I.p[i]) < I.errpn) acc[1] += 1.0;
        }
    }
    write_partials<2>(acc, 0, ro, gstart, it);
}

// snorm expand step (src/m_snorm.f90:253-286): returns (#moved in, #small dd)
__global__ void __launch_bounds__(ELEM_THREADS) k_sn_expand(const NcgItem* items, const int* gstart,
                                                            int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] == EXTER && I.h[i] - I.pen < 0.0) {
            double d = I.h[i] - I.pen + I.unn[i];
            if (d < -I.tol) {
                I.el[i] = ADHES;
                acc[0] += 1.0;
            }
            if (fabs(d) < I.tol) acc[1] += 1.0;
        }
    }
    write_partials<2>(acc, 0, ro, gstart, it);
}

// final regularisation (src/m_snorm.f90:315-325)
__global__ void __launch_bounds__(ELEM_THREADS) k_sn_final(const NcgItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const NcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.p[i] < 0.0 && I.el[i] >= ADHES) {
            I.p[i] = 0.0;
            I.el[i] = EXTER;
            acc[0] += 1.0;
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// gf3_AijPj with unit tractions at one element: sum over rows jy of the coefficients for
// jx in [row1st-1, rowlst+1] of the current contact area (src/m_aijpj.f90:158-185,
// src/m_gridfunc.f90:582-637).  One group per item, one thread per row.
__global__ void __launch_bounds__(ROW_THREADS) k_aij_ones(const AijItem* items, const int* gstart,
                                                          int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const AijItem& I = items[it];
    double acc[1] = {0.0};
    int rows_per_group = ROW_THREADS;
    int jy0 = (blockIdx.x - gstart[it]) * rows_per_group;
    int jy = jy0 + threadIdx.x + 1;   // 1-based row
    if (jy <= I.my) {
        const int8_t* row = I.el + (size_t)(jy - 1) * I.mx;
        int first = 1;
        while (row[first - 1] <= EXTER && first != I.mx) first++;
        int last = 0;
        for (int ix = first; ix <= I.mx; ix++)
            if (row[ix - 1] >= ADHES) last = ix;
        int j0 = first - 1 < 1 ? 1 : first - 1;
        int j1 = last + 1 > I.mx ? I.mx : last + 1;
        double rowsum = 0.0;
        for (int jx = j0; jx <= j1; jx++) {
            int dx = I.ix - jx, dy = I.iy - jy;
            if (dx < -I.cf_mx || dx >= I.cf_mx || dy < -I.cf_my || dy >= I.cf_my) continue;
            rowsum += I.cf[(size_t)(dy + I.cf_my) * (2 * I.cf_mx) + (dx + I.cf_mx)];
        }
        acc[0] = rowsum;
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// ------------------------------------------------------------------------------------------
// TangCG element steps (src/m_solvpt.f90:1948-2384)
// ------------------------------------------------------------------------------------------

__device__ inline void nt_from_p(double px, double py, double& nx, double& ny)
{
    double th = atan2(py, px);
    nx = cos(th);
    ny = sin(th);
}

__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_nt(const TcgItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double nx, ny;
        nt_from_p(I.px[i], I.py[i], nx, ny);
        I.nx[i] = nx;
        I.ny[i] = ny;
    }
}

// r = -s, projected on t in the slip area
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_rset(const TcgItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double rx = -I.sx[i], ry = -I.sy[i];
        if (I.el[i] == SLIP) {
            double tx = -I.ny[i], ty = I.nx[i];
            double perp = tx * rx + ty * ry;
            rx = perp * tx;
            ry = perp * ty;
        }
        I.rx[i] = rx;
        I.ry[i] = ry;
    }
}

// Jacobian diagonal scaling and slip projection of z; returns (z.q, v.q) over all elements
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_zprep(const TcgItem* items, const int* gstart,
                                                            int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double zx = I.zx[i], zy = I.zy[i];
        if (I.el[i] == SLIP) {
            double g = I.mu * I.pn[i];
            double snrm = -I.sx[i] * I.nx[i] - I.sy[i] * I.ny[i];
            zx = zx / (I.c11 + I.ga * snrm / g);
            zy = zy / (I.c22 + I.ga * snrm / g);
            double tx = -I.ny[i], ty = I.nx[i];
            double perp = tx * zx + ty * zy;
            zx = perp * tx;
            zy = perp * ty;
        } else {
            zx = zx / I.c11;
            zy = zy / I.c22;
        }
        I.zx[i] = zx;
        I.zy[i] = zy;
        acc[0] += zx * I.qx[i] + zy * I.qy[i];
        acc[1] += I.vx[i] * I.qx[i] + I.vy[i] * I.qy[i];
    }
    write_partials<2>(acc, 0, ro, gstart, it, I.sc);
}

// v = z + beta v with beta from zprep's (z.q, v.q) in sc[0..1] (src/m_solvpt.f90:2150-2160)
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_vupd(const TcgItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    (void)ro;
    const double small = 1e-6, zq = I.sc[0], vq0 = I.sc[1];
    double beta;
    if (fabs(zq) < 1e-60 || fabs(vq0) < small * fabs(zq)) beta = 0.0;
    else beta = -zq / vq0;
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        I.vx[i] = 1.0 * I.zx[i] + beta * I.vx[i];
        I.vy[i] = 1.0 * I.zy[i] + beta * I.vy[i];
    }
}

// q: slip projection and curvature term; returns (r.v, v.q)
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_qpost(const TcgItem* items, const int* gstart,
                                                            int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double qx = I.qx[i], qy = I.qy[i];
        if (I.el[i] == SLIP) {
            double tx = -I.ny[i], ty = I.nx[i];
            double perp = tx * qx + ty * qy;
            qx = perp * tx;
            qy = perp * ty;
            double g = I.mu * I.pn[i];
            double snrm = (I.nx[i] * I.sx[i] + I.ny[i] * I.sy[i]) / g;
            qx -= snrm * I.vx[i];
            qy -= snrm * I.vy[i];
        }
        I.qx[i] = qx;
        I.qy[i] = qy;
        acc[0] += I.rx[i] * I.vx[i] + I.ry[i] * I.vy[i];
        acc[1] += I.vx[i] * qx + I.vy[i] * qy;
    }
    write_partials<2>(acc, 0, ro, gstart, it, I.sc ? I.sc + 2 : nullptr);
}

// pold = p; p += alpha v (interior); slip clipped to the bound; returns (sum v^2, sum p^2), or
// sum (p - p0)^2 in a sensitivity solve (src/m_solvpt.f90:2209-2216), then (r.v, v.q).  With
// sc set, alpha is formed here from qpost's (r.v, v.q) in sc[2..3] (:2190-2200).
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_pupd(const TcgItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double alpha = I.alpha, rv = 0.0, vq = 0.0;
    if (I.sc) {
        const double small = 1e-6;
        rv = I.sc[2];
        vq = I.sc[3];
        if (fabs(rv) < 1e-60) alpha = 0.0;
        else if (fabs(vq) < small * fabs(rv)) alpha = 1.0;
        else alpha = rv / vq;
    }
    // (r.v, v.q) enter the sums once, from the item's first group's first thread
    const bool first = blockIdx.x == gstart[it] && threadIdx.x == 0;
    double acc[4] = {0.0, 0.0, first ? rv : 0.0, first ? vq : 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double px = I.px[i], py = I.py[i];
        I.poldx[i] = px;
        I.poldy[i] = py;
        int8_t s = I.el[i];
        if (s >= ADHES) {
            px += alpha * I.vx[i];
            py += alpha * I.vy[i];
        }
        if (s == SLIP) {
            double g = I.mu * I.pn[i];
            double ptabs = sqrt(px * px + py * py);
            px = px * g / ptabs;
            py = py * g / ptabs;
        }
        I.px[i] = px;
        I.py[i] = py;
        acc[0] += I.vx[i] * I.vx[i] + I.vy[i] * I.vy[i];
        if (I.p0x) {
            px = px + -1.0 * I.p0x[i];
            py = py + -1.0 * I.p0y[i];
        }
        acc[1] += px * px + py * py;
    }
    write_partials<4>(acc, 0, ro, gstart, it);
}

// adhesion elements above the traction bound move to slip; returns #moved
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_as(const TcgItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] == ADHES) {
            double g = I.mu * I.pn[i];
            double px = I.px[i], py = I.py[i];
            double ptabs = sqrt(px * px + py * py);
            if (ptabs > g) {
                I.el[i] = SLIP;
                I.px[i] = px * g / ptabs;
                I.py[i] = py * g / ptabs;
                acc[0] += 1.0;
            }
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// slip elements with s.p > 0 move to adhesion; pold -= p; returns (#moved, sum (pold-p)^2)
__global__ void __launch_bounds__(ELEM_THREADS) k_tcg_sa(const TcgItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const TcgItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] == SLIP) {
            double snrm = I.sx[i] * I.px[i] + I.sy[i] * I.py[i];
            if (snrm > 0.0) {
                I.el[i] = ADHES;
                acc[0] += 1.0;
            }
        }
        double dx = I.poldx[i] - I.px[i], dy = I.poldy[i] - I.py[i];
        I.poldx[i] = dx;
        I.poldy[i] = dy;
        acc[1] += dx * dx + dy * dy;
    }
    write_partials<2>(acc, 0, ro, gstart, it);
}

// ------------------------------------------------------------------------------------------
// stang outer loop moves (src/m_stang.f90:433-508)
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(ELEM_THREADS) k_st_as(const StItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const StItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] == ADHES) {
            double px = I.px[i], py = I.py[i];
            double pabs = sqrt(px * px + py * py);
            double g = I.mu * I.pn[i];
            if (pabs >= g + I.tol) {
                I.el[i] = SLIP;
                I.px[i] = px * g / pabs;
                I.py[i] = py * g / pabs;
                acc[0] += 1.0;
            }
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

__global__ void __launch_bounds__(ELEM_THREADS) k_st_sa(const StItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const StItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.el[i] == SLIP) {
            double ww = I.sx[i] * I.px[i] + I.sy[i] * I.py[i];
            double tl = I.tol1 * fabs(I.px[i]) + I.tol2 * fabs(I.py[i]) +
                        I.errpt * (fabs(I.sx[i]) + fabs(I.sy[i]));
            if (ww > tl) {
                I.el[i] = ADHES;
                acc[0] += 1.0;
            }
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// ------------------------------------------------------------------------------------------
// GDsteady (src/gdsteady.f90)
// ------------------------------------------------------------------------------------------

// The GD row recurrences (project_searchdir, apply_trcbnd) run from the leading edge
// (ix = mx) to the trailing edge along each row.  A work-group takes a tile of whole rows:
// it loads the recurrence inputs with coalesced reads into LDS, one thread per row runs the
// recurrence in LDS, and the element-wise rest is done by all threads with coalesced access.
constexpr int GD_THREADS = 256;
constexpr int GD_TILE = 2048;          // elements per tile (probe, update)
constexpr int GD_TILE_PROJ = 1024;     // elements per tile (project)

__host__ __device__ inline int gd_rows_per_tile(int mx, int my, int tile)
{
    int r = tile / mx;
    if (r < 1) r = 1;
    if (r > GD_THREADS) r = GD_THREADS;
    return r < my ? r : my;
}
__host__ __device__ inline int gd_row_stride(int mx) { return mx | 1; }

// compute_dp, src/gdsteady.f90:610-635: dp = backward differences along x (element-wise)
__global__ void __launch_bounds__(ELEM_THREADS) k_gd_dp(const GdItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int ix = i % I.mx;
        if (ix == I.mx - 1) {
            I.dpx[i] = I.px[i];
            I.dpy[i] = I.py[i];
        } else {
            I.dpx[i] = I.px[i] - I.px[i + 1];
            I.dpy[i] = I.py[i] - I.py[i + 1];
        }
    }
}

// project_searchdir, src/gdsteady.f90:639-803: dv from the residual r, with v // t in slip
__global__ void __launch_bounds__(GD_THREADS) k_gd_project(const GdItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    extern __shared__ double gd_lds[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    const int mx = I.mx, S = gd_row_stride(mx);
    const int R = gd_rows_per_tile(mx, I.my, GD_TILE_PROJ);
    const int row0 = (blockIdx.x - gstart[it]) * R;
    const int nrow = min(R, I.my - row0);
    const int ne = nrow * mx;
    const size_t g0 = (size_t)row0 * mx;
    double* lrx = gd_lds;
    double* lry = lrx + R * S;
    double* lg = lry + R * S;
    double* lnx = lg + R * S;        // n, then v
    double* lny = lnx + R * S;
    double* ldx = lny + R * S;       // dv
    double* ldy = ldx + R * S;
    int8_t* lel = (int8_t*)(ldy + R * S);
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, l = r * S + (e - r * mx);
        size_t i = g0 + e;
        int8_t st = I.el[i];
        lel[l] = st;
        lrx[l] = I.rx[i];
        lry[l] = I.ry[i];
        if (st == SLIP) {
            lg[l] = I.mu * I.pn[i];
            lnx[l] = I.nx[i];
            lny[l] = I.ny[i];
        }
    }
    __syncthreads();
    if (threadIdx.x < nrow) {
        const int l0 = threadIdx.x * S;
        const int imeth = I.imeth, kdown = I.kdown;
        double vxn = 0.0, vyn = 0.0;   // v at ix+1
        lnx[l0 + mx - 1] = 0.0;
        lny[l0 + mx - 1] = 0.0;
        ldx[l0 + mx - 1] = lrx[l0 + mx - 1];
        ldy[l0 + mx - 1] = lry[l0 + mx - 1];
        int last_se = 1 << 30;   // smallest processed (1-based) ix in slip/exterior, E_down
        for (int ix = mx - 2; ix >= 0; ix--) {
            const int l = l0 + ix;
            int8_t st = lel[l];
            double dvx = lrx[l], dvy = lry[l], vx, vy;
            if (st == ADHES) {
                if (imeth == 1) {
                    vx = vxn + dvx;
                    vy = vyn + dvy;
                } else if (imeth == 3) {
                    vx = I.fdecay * vxn + dvx;
                    vy = I.fdecay * vyn + dvy;
                    dvx = vx - vxn;
                    dvy = vy - vyn;
                } else {
                    int ixf = ix + 1;
                    if (ixf + kdown <= mx) {
                        double inx = lrx[l + kdown], iny = lry[l + kdown];
                        if (last_se <= ixf + kdown - 1) { inx = 0.0; iny = 0.0; }
                        vx = vxn + dvx - inx;
                        vy = vyn + dvy - iny;
                    } else {
                        vx = vxn + dvx;
                        vy = vyn + dvy;
                    }
                    dvx = vx - vxn;
                    dvy = vy - vyn;
                }
            } else if (st == SLIP) {
                double g = lg[l];
                double tx = -lny[l], ty = lnx[l];
                double vt = tx * dvx + ty * dvy;
                vt = copysign(1.0, vt) * fmin(fabs(vt), I.fac_v * g);
                vx = tx * vt;
                vy = ty * vt;
                dvx = vx - vxn;
                dvy = vy - vyn;
                last_se = ix + 1;
            } else {
                vx = 0.0;
                vy = 0.0;
                dvx = -vxn;
                dvy = -vyn;
                last_se = ix + 1;
            }
            lnx[l] = vx;
            lny[l] = vy;
            ldx[l] = dvx;
            ldy[l] = dvy;
            vxn = vx;
            vyn = vy;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, l = r * S + (e - r * mx);
        size_t i = g0 + e;
        I.vx[i] = lnx[l];
        I.vy[i] = lny[l];
        I.dvx[i] = ldx[l];
        I.dvy[i] = ldy[l];
    }
}

// line-search probe at step alpha (src/gdsteady.f90:1090-1158): apply_trcbnd on
// dp + alpha dv, then the residual terms of the new orientation.  Returns
// (rho, c0adh, c1adh, c0slp, c1slp, c2slp) and, when I.first, also
// (sum r^2, sum r.q dscl, sum_int (dscl q)^2, sum_int r^2).
__global__ void __launch_bounds__(GD_THREADS) k_gd_probe(const GdItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    extern __shared__ double gd_lds[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    const int mx = I.mx, S = gd_row_stride(mx);
    const int R = gd_rows_per_tile(mx, I.my, GD_TILE);
    const int row0 = (blockIdx.x - gstart[it]) * R;
    const int nrow = min(R, I.my - row0);
    const int ne = nrow * mx;
    const size_t g0 = (size_t)row0 * mx;
    const double a = I.alpha;
    double* lpx = gd_lds;            // dp + alpha dv, then the new p
    double* lpy = lpx + R * S;
    double* lg = lpy + R * S;
    int8_t* lel = (int8_t*)(lg + R * S);
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, l = r * S + (e - r * mx);
        size_t i = g0 + e;
        int8_t st = I.el[i];
        lel[l] = st;
        if (st >= ADHES) {
            lg[l] = I.mu * I.pn[i];
            lpx[l] = I.dpx[i] + a * I.dvx[i];
            lpy[l] = I.dpy[i] + a * I.dvy[i];
        }
    }
    __syncthreads();
    if (threadIdx.x < nrow) {
        const int l0 = threadIdx.x * S;
        double pxn = 0.0, pyn = 0.0;   // new p at ix+1
        for (int ix = mx - 1; ix >= 0; ix--) {
            const int l = l0 + ix;
            int8_t st = lel[l];
            double px, py;
            if (ix == mx - 1 || st <= EXTER) {
                px = 0.0;
                py = 0.0;
            } else {
                double g = lg[l];
                px = pxn + lpx[l];
                py = pyn + lpy[l];
                double ptabs = sqrt(px * px + py * py);
                if (st == SLIP || ptabs >= g) {
                    px = px * g / ptabs;
                    py = py * g / ptabs;
                }
            }
            lpx[l] = px;
            lpy[l] = py;
            pxn = px;
            pyn = py;
        }
    }
    __syncthreads();
    double acc[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double tiny = 1e-12;
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, l = r * S + (e - r * mx);
        size_t ii = g0 + e;
        int8_t st = lel[l];
        double dscl = I.dscl[ii];
        double sx = I.sx[ii], sy = I.sy[ii], qx = I.qx[ii], qy = I.qy[ii];
        if (I.first) {
            double rx = I.rx[ii], ry = I.ry[ii];
            acc[6] += rx * rx + ry * ry;
            acc[7] += (rx * qx + ry * qy) * dscl;
            if (st >= ADHES) {
                acc[8] += dscl * dscl * (qx * qx + qy * qy);
                acc[9] += rx * rx + ry * ry;
            }
        }
        if (st != ADHES && st != SLIP) continue;
        double px = lpx[l], py = lpy[l], g = lg[l];
        double txo = -I.ny[ii], tyo = I.nx[ii];
        double vt = I.vx[ii] * txo + I.vy[ii] * tyo;
        double dth_da = (g * vt) / (fmax(tiny, g * g) + a * a * vt * vt);
        double nxn, nyn;
        nt_from_p(px, py, nxn, nyn);
        double txn = -nyn, tyn = nxn;
        double sn = sx * nxn + sy * nyn;
        int8_t elnew = st;
        if (elnew == SLIP && sn > 0.0) elnew = ADHES;
        double stt = sx * txn + sy * tyn, rt = -stt;
        double qn = qx * nxn + qy * nyn, qt = qx * txn + qy * tyn;
        double d2 = dscl * dscl;
        if (elnew == ADHES) {
            acc[1] += 2.0 * d2 * (sx * qx + sy * qy);
            acc[2] += 2.0 * d2 * (qx * qx + qy * qy);
            double ex = sx + a * qx, ey = sy + a * qy;
            acc[0] += d2 * (ex * ex + ey * ey);
        } else {
            acc[3] -= 2.0 * d2 * (rt * qt - rt * sn * dth_da);
            acc[4] -= 2.0 * d2 * (-qt * qt + qt * sn * dth_da - rt * qn * dth_da);
            acc[5] -= 2.0 * d2 * (qt * qn * dth_da);
            double ex = stt + a * qt;
            acc[0] += d2 * ex * ex;
        }
    }
    write_partials<10>(acc, 0, ro, gstart, it);
}

// pold = p; dp += alpha dv on the interior; apply_trcbnd (src/gdsteady.f90:807-847); n from p
__global__ void __launch_bounds__(GD_THREADS) k_gd_update(const GdItem* items, const int* gstart,
                                                          int nitems, RedOut ro)
{
    extern __shared__ double gd_lds[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    const int mx = I.mx, S = gd_row_stride(mx);
    const int R = gd_rows_per_tile(mx, I.my, GD_TILE);
    const int row0 = (blockIdx.x - gstart[it]) * R;
    const int nrow = min(R, I.my - row0);
    const int ne = nrow * mx;
    const size_t g0 = (size_t)row0 * mx;
    double* lpx = gd_lds;            // dp + alpha dv, then the new p
    double* lpy = lpx + R * S;
    double* lg = lpy + R * S;
    int8_t* lel = (int8_t*)(lg + R * S);
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, l = r * S + (e - r * mx);
        size_t i = g0 + e;
        int8_t st = I.el[i];
        lel[l] = st;
        lg[l] = I.mu * I.pn[i];
        double dpx = I.dpx[i], dpy = I.dpy[i];
        if (st >= ADHES) {
            dpx += I.alpha * I.dvx[i];
            dpy += I.alpha * I.dvy[i];
        }
        lpx[l] = dpx;
        lpy[l] = dpy;
        I.poldx[i] = I.px[i];
        I.poldy[i] = I.py[i];
        // the leading-edge column keeps dp (not recomputed from p)
        if (e - r * mx == mx - 1) {
            I.dpx[i] = dpx;
            I.dpy[i] = dpy;
        }
    }
    __syncthreads();
    if (threadIdx.x < nrow) {
        const int l0 = threadIdx.x * S;
        double pxn = 0.0, pyn = 0.0;
        for (int ix = mx - 1; ix >= 0; ix--) {
            const int l = l0 + ix;
            int8_t st = lel[l];
            double px, py;
            if (ix == mx - 1 || st <= EXTER) {
                px = 0.0;
                py = 0.0;
            } else {
                double g = lg[l];
                px = pxn + lpx[l];
                py = pyn + lpy[l];
                double ptabs = sqrt(px * px + py * py);
                if (st == SLIP || ptabs >= g) {
                    px = px * g / ptabs;
                    py = py * g / ptabs;
                }
            }
            lpx[l] = px;
            lpy[l] = py;
            pxn = px;
            pyn = py;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += GD_THREADS) {
        int r = e / mx, ix = e - r * mx, l = r * S + ix;
        size_t ii = g0 + e;
        double px = lpx[l], py = lpy[l];
        if (ix != mx - 1) {
            I.dpx[ii] = px - lpx[l + 1];
            I.dpy[ii] = py - lpy[l + 1];
        }
        I.px[ii] = px;
        I.py[ii] = py;
        double nx, ny;
        nt_from_p(px, py, nx, ny);
        I.nx[ii] = nx;
        I.ny[ii] = ny;
    }
}

// 1/d and 1/sqrt(x) on the latency path of SteadyGS: the hardware estimates refined by two
// Newton steps each (within an ulp or two of the correctly rounded values, without the IEEE
// division and square-root sequences)
__device__ inline double fast_rcp(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}
__device__ inline double fast_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double t = fma(-h * y, y, 0.5);   // y <- y + y (1/2 - x y^2 / 2)
    y = fma(y, t, y);
    t = fma(-h * y, y, 0.5);
    return fma(y, t, y);
}

// plstrc without plasticity (elmtrc), src/m_solvpt.f90:3258-3807; the plastic branch only
// reports a violated constraint when tau_c0 = 1e20 and k_tau = 0
// scale_dir: with omegas = 1 put the slip traction on the bound along the Newton solution
// directly instead of through atan2, cos and sin (the same direction up to rounding), divide
// by refined reciprocals and reciprocal square roots, and test the adhesion bound on squares
// (the same values up to rounding); used on the sequential path of SteadyGS, where the chain
// of FP64 divisions, square roots and transcendentals is the latency of every element
__device__ void elmtrc(int8_t& el, const double c11, const double c12, const double c21,
                       const double c22, double eps, double pr[3], double mus, double s[2],
                       double omegah = 1.0, double omegas = 1.0, bool scale_dir = false)
{
    // (a / d, b / d), or through 1 / d on the latency path
    auto pair_div = [scale_dir](double a, double b, double d, double& qa, double& qb) {
        if (scale_dir) {
            const double r = fast_rcp(d);
            qa = a * r;
            qb = b * r;
        } else {
            qa = a / d;
            qb = b / d;
        }
    };
    const int maxnr = 10;
    const double prc = 0.0001, taucs = 1e20, PI = 3.14159265358979323846;
    double pold1 = pr[0], pold2 = pr[1];
    double si0x = s[0] - c11 * pr[0] - c12 * pr[1];
    double si0y = s[1] - c21 * pr[0] - c22 * pr[1];
    double six = 0.0, siy = 0.0;
    int itry = 0, violat = 1;
    while (itry <= 3 && violat == 1) {
        itry++;
        if (el == ADHES && scale_dir) {
            // the same steps with the bound compared on squares and the reciprocals refined
            const double trc_bound = mus * pr[2];
            const double detinv = fast_rcp(c11 * c22 - c12 * c21);
            const double dp0 = (-c22 * s[0] + c12 * s[1]) * detinv;
            const double dp1 = (c21 * s[0] - c11 * s[1]) * detinv;
            pr[0] = pold1 + omegah * dp0;
            pr[1] = pold2 + omegah * dp1;
            const double p2 = pr[0] * pr[0] + pr[1] * pr[1];
            const double bnd = fmin(trc_bound, taucs);
            if (p2 <= bnd * bnd) {
                violat = 0;
            } else {
                const double q0 = pold1 + dp0, q1 = pold2 + dp1;
                if (q0 * q0 + q1 * q1 <= bnd * bnd) {
                    const double f = bnd * fast_rsqrt(p2);
                    pr[0] *= f;
                    pr[1] *= f;
                    violat = 0;
                } else {
                    violat = 1;
                }
            }
            six = si0x + c11 * pr[0] + c12 * pr[1];
            siy = si0y + c21 * pr[0] + c22 * pr[1];
        } else if (el == ADHES) {
            double trc_bound = mus * pr[2];
            double detinv = 1.0 / (c11 * c22 - c12 * c21);
            double dp0 = (-c22 * s[0] + c12 * s[1]) * detinv;
            double dp1 = (c21 * s[0] - c11 * s[1]) * detinv;
            pr[0] = pold1 + omegah * dp0;
            pr[1] = pold2 + omegah * dp1;
            double pabs2 = sqrt(pr[0] * pr[0] + pr[1] * pr[1]);
            double bnd = fmin(trc_bound, taucs);
            if (pabs2 <= bnd) {
                violat = 0;
            } else {
                double pabs1 = sqrt((pold1 + dp0) * (pold1 + dp0) + (pold2 + dp1) * (pold2 + dp1));
                if (pabs1 <= bnd) {
                    pair_div(pr[0] * bnd, pr[1] * bnd, pabs2, pr[0], pr[1]);
                    violat = 0;
                } else {
                    violat = 1;
                }
            }
            six = si0x + c11 * pr[0] + c12 * pr[1];
            siy = si0y + c21 * pr[0] + c22 * pr[1];
        } else if (el == SLIP) {
            violat = 0;
            double trc_bound = mus * pr[2];
            if (scale_dir) {
                const double f = trc_bound * fast_rsqrt(si0x * si0x + si0y * si0y);
                pr[0] = -f * si0x;
                pr[1] = -f * si0y;
            } else {
                double sabs = sqrt(si0x * si0x + si0y * si0y);
                pair_div(-trc_bound * si0x, -trc_bound * si0y, sabs, pr[0], pr[1]);
            }
            six = si0x + c11 * pr[0] + c12 * pr[1];
            siy = si0y + c21 * pr[0] + c22 * pr[1];
            double fk0 = pr[1] * six - pr[0] * siy;
            double fk1 = pr[0] * pr[0] + pr[1] * pr[1] - trc_bound * trc_bound;
            int itnr = 0;
            while (itnr == 0 || (itnr < maxnr && fabs(fk0) + fabs(fk1) >= prc * eps * trc_bound)) {
                itnr++;
                double g00 = -2.0 * c21 * pr[0] - si0y + (c11 - c22) * pr[1];
                double g01 = 2.0 * c12 * pr[1] + si0x + (c11 - c22) * pr[0];
                double g10 = 2.0 * pr[0];
                double g11 = 2.0 * pr[1];
                double det = g00 * g11 - g01 * g10;
                if (det == 0.0) {
                    itnr = maxnr;
                } else {
                    double d0, d1;
                    pair_div(-(g11 * fk0 - g01 * fk1), -(-g10 * fk0 + g00 * fk1), det, d0, d1);
                    pr[0] += d0;
                    pr[1] += d1;
                }
                six = si0x + c11 * pr[0] + c12 * pr[1];
                siy = si0y + c21 * pr[0] + c22 * pr[1];
                fk0 = pr[1] * six - pr[0] * siy;
                fk1 = pr[0] * pr[0] + pr[1] * pr[1] - trc_bound * trc_bound;
            }
            if (scale_dir && omegas == 1.0) {
                // alph0 + (alph1 - alph0) is alph1 up to rounding (the 2 pi wrap keeps the
                // direction): the traction at the bound in the Newton solution's direction
                const double p2 = pr[0] * pr[0] + pr[1] * pr[1];
                if (p2 > 0.0) {
                    const double f = trc_bound * fast_rsqrt(p2);
                    pr[0] *= f;
                    pr[1] *= f;
                } else {
                    pr[0] = trc_bound;
                    pr[1] = 0.0;
                }
            } else {
                double alph0 = atan2(pold2, pold1), alph1 = atan2(pr[1], pr[0]);
                double dalph = alph1 - alph0;
                if (dalph < -PI) dalph += 2.0 * PI;
                if (dalph > PI) dalph -= 2.0 * PI;
                alph1 = alph0 + omegas * dalph;
                pr[0] = trc_bound * cos(alph1);
                pr[1] = trc_bound * sin(alph1);
            }
            if (fabs(pr[0]) > fabs(pr[1])) {
                if (pr[0] * six > 0.0) violat = 1;
            } else {
                if (pr[1] * siy > 0.0) violat = 1;
            }
        } else {
            violat = 1;
        }
        if (itry <= 3 && violat == 1) {
            if (el == ADHES) el = SLIP;
            else if (el == SLIP) el = PLAST;
            else el = ADHES;
        }
    }
    s[0] = six;
    s[1] = siy;
}

// element-state moves after a GD step (src/gdsteady.f90:375-418); returns #changes
__global__ void __launch_bounds__(ELEM_THREADS) k_gd_moves(const GdItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[1] = {0.0};
    const double tiny = 1e-12;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int8_t s = I.el[i];
        double g = I.mu * I.pn[i];
        if (s == ADHES) {
            double ptabs = sqrt(I.px[i] * I.px[i] + I.py[i] * I.py[i]);
            if (ptabs >= g - tiny) {
                I.el[i] = SLIP;
                acc[0] += 1.0;
            }
        } else if (s == SLIP) {
            double pr[3] = {I.px[i], I.py[i], I.pn[i]};
            double si[2] = {I.sx[i], I.sy[i]};
            int8_t elopt = s;
            elmtrc(elopt, I.c11, I.c12, I.c21, I.c22, 1e-6, pr, I.mu, si);
            double snrm = I.nx[i] * I.sx[i] + I.ny[i] * I.sy[i];
            double cosdth = (I.nx[i] * pr[0] + I.ny[i] * pr[1]) / g;
            if ((snrm > 0.0 && (I.dscl[i] >= 0.001 || elopt == ADHES)) || cosdth <= -0.5) {
                I.el[i] = ADHES;
                acc[0] += 1.0;
            }
        }
    }
    write_partials<1>(acc, 0, ro, gstart, it);
}

// angle-split of a slip element (compute_diagscaling, src/gdsteady.f90:927-980); returns
// false when the angle error stays below tiny_err (elnew/fac_s then keep earlier values)
__device__ inline bool gd_slip_split(const GdItem& I, int ii, int8_t& elnew, double& fac_s)
{
    const double tiny_err = 1e-6, PI = 3.14159265358979323846;
    double pr[3] = {I.px[ii], I.py[ii], I.pn[ii]};
    double si[2] = {I.sx[ii], I.sy[ii]};
    double th_p0 = atan2(pr[1], pr[0]);
    double th_s0 = atan2(-si[1], -si[0]);
    double err_0 = th_s0 - th_p0;
    if (fabs(err_0) >= PI) err_0 = err_0 - rint(err_0 / (2.0 * PI)) * 2.0 * PI;
    if (fabs(err_0) <= tiny_err) {
        pr[0] = I.mu * pr[2] * cos(th_p0 + 0.1);
        pr[1] = I.mu * pr[2] * sin(th_p0 + 0.1);
        si[0] = I.sx[ii] + I.c11 * (pr[0] - I.px[ii]) + I.c12 * (pr[1] - I.py[ii]);
        si[1] = I.sy[ii] + I.c21 * (pr[0] - I.px[ii]) + I.c22 * (pr[1] - I.py[ii]);
        th_p0 = atan2(pr[1], pr[0]);
        th_s0 = atan2(-si[1], -si[0]);
        err_0 = th_s0 - th_p0;
        if (fabs(err_0) >= PI) err_0 = err_0 - rint(err_0 / (2.0 * PI)) * 2.0 * PI;
    }
    if (fabs(err_0) <= tiny_err) return false;
    int8_t en = SLIP;
    elmtrc(en, I.c11, I.c12, I.c21, I.c22, 1e-6, pr, I.mu, si);
    elnew = en;
    if (en == SLIP) {
        double th_s1 = atan2(-si[1], -si[0]);
        double upd_s = th_s1 - th_s0;
        if (fabs(upd_s) >= PI) upd_s = upd_s - rint(upd_s / (2.0 * PI)) * 2.0 * PI;
        fac_s = fabs(upd_s) / fabs(err_0);
    }
    return true;
}

// compute_diagscaling and the scaled residual (src/gdsteady.f90:901-1042, 429-439); when
// not I.first also pold -= p on the interior.  Returns (sum (pold)^2, sum p^2, sum_int r^2).
__global__ void __launch_bounds__(ELEM_THREADS) k_gd_diagres(const GdItem* items, const int* gstart,
                                                             int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const GdItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int8_t s = I.el[i];
        int8_t elnew = EXTER;
        double fac_s = 0.0;
        bool have = false;
        if (s == SLIP) have = gd_slip_split(I, i, elnew, fac_s);
        if (s == SLIP && !have) {
            // the reference keeps elnew/fac_s of the last slip element it split
            for (int j = i - 1; j >= 0; j--) {
                if (I.el[j] != SLIP) continue;
                if (gd_slip_split(I, j, elnew, fac_s)) break;
            }
        }
        double dscl = I.dscl[i];
        if (s == SLIP && elnew == SLIP) {
            dscl = I.d_slp * pow(fac_s, I.pow_s);
        } else if (s == ADHES || (s == SLIP && elnew == ADHES)) {
            int ix = i % I.mx + 1;
            int jx = ix;
            while (jx > 1 && I.el[i - ix + jx - 1] == ADHES) jx--;
            double dnew;
            if (I.d_cns < I.d_ifc)
                dnew = fmax(I.d_cns, I.d_ifc + (ix - jx - 1) * fmin(0.0, I.d_lin));
            else
                dnew = fmin(I.d_cns, I.d_ifc + (ix - jx - 1) * fmax(0.0, I.d_lin));
            dscl = fmin(50.0 * dscl, dnew);
        } else if (s != ADHES) {
            dscl = 0.0;
        }
        I.dscl[i] = dscl;
        double rx = 0.0, ry = 0.0;
        if (s == ADHES) {
            rx = -I.sx[i] * dscl;
            ry = -I.sy[i] * dscl;
        } else if (s == SLIP) {
            double tx = -I.ny[i], ty = I.nx[i];
            double st = tx * I.sx[i] + ty * I.sy[i];
            rx = -st * tx * dscl;
            ry = -st * ty * dscl;
        }
        I.rx[i] = rx;
        I.ry[i] = ry;
        double px = I.px[i], py = I.py[i];
        if (!I.first) {
            double ox = I.poldx[i], oy = I.poldy[i];
            if (s >= ADHES) {
                ox -= px;
                oy -= py;
                I.poldx[i] = ox;
                I.poldy[i] = oy;
            }
            acc[0] += ox * ox + oy * oy;
        }
        acc[1] += px * px + py * py;
        if (s >= ADHES) acc[2] += rx * rx + ry * ry;
    }
    write_partials<3>(acc, 0, ro, gstart, it);
}

// ------------------------------------------------------------------------------------------
// FFT products (src/m_aijpj.f90:712-1015)
// ------------------------------------------------------------------------------------------

template <typename R>
struct LdsSync {
    __device__ void operator()() const { __syncthreads(); }
};

// rows per work-group for a complex row length M, bounded by the LDS budget
__host__ __device__ inline int rows_per_group(int M) { int r = 2048 / M; return r < 1 ? 1 : (r > 32 ? 32 : r); }
// columns per work-group for a column length L and nc components
__host__ __device__ inline int cols_per_group(int L, int nc) { int r = 2048 / (L * nc); return r < 1 ? 1 : (r > 16 ? 16 : r); }
// the shapes of one product item: its own (chosen by the engine's tuner) or the defaults
__host__ __device__ inline int conv_rows(const ConvItem& c) { return c.rpg > 0 ? c.rpg : rows_per_group(c.Lx / 2); }
__host__ __device__ inline int conv_cols(const ConvItem& c, int nc) { return c.cpg > 0 ? c.cpg : cols_per_group(c.Ly, nc); }

template <typename R>
__device__ inline R to_r(double v) { return (R)v; }

// forward real row transforms: rows y < ny_rows of `in` (row length nxr, zero padded to Lx)
// -> S[c][y*H + kx]
template <typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_rowfwd(const ConvItem* items, const int* gstart,
                                                        int nitems)
{
    extern __shared__ char lds_raw[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const ConvItem& I = items[it];
    const int Lx = I.Lx, M = Lx / 2, H = M + 1;
    // mode 0: inputs are grid functions (rows my, length mx); modes 1/2: dense Lx*Ly arrays
    const int nrow = I.mode == 0 ? I.my : I.Ly;
    const int nxr = I.mode == 0 ? I.mx : I.Lx;
    const int rpg = conv_rows(I);
    const int ngy = (nrow + rpg - 1) / rpg;
    int t = blockIdx.x - gstart[it];
    int c = t / ngy;
    int y0 = (t - c * ngy) * rpg;
    int nr = min(rpg, nrow - y0);
    cpx<R>* a = (cpx<R>*)lds_raw;
    cpx<R>* b = a + rpg * M;
    const double* src = I.in[c];
    for (int idx = threadIdx.x; idx < nr * M; idx += blockDim.x) {
        int r = idx / M, n = idx - r * M;
        int x0 = 2 * n;
        const double* row = src + (size_t)(y0 + r) * nxr;
        R re = x0 < nxr ? to_r<R>(row[x0]) : (R)0;
        R im = x0 + 1 < nxr ? to_r<R>(row[x0 + 1]) : (R)0;
        a[idx] = {re, im};
    }
    __syncthreads();
    FftPlan plan = make_plan(M);
    cpx<R>* res = fft_run<R>(a, b, plan, nr, (R)-1, (const cpx<R>*)I.twM, threadIdx.x, blockDim.x,
                             LdsSync<R>());
    cpx<R>* S = (cpx<R>*)I.wsS + (size_t)c * nrow * H;
    const cpx<R>* twL = (const cpx<R>*)I.twLx;
    for (int idx = threadIdx.x; idx < nr * H; idx += blockDim.x) {
        int r = idx / H, k = idx - r * H;
        cpx<R> Z = res[r * M + (k % M)];
        cpx<R> Zc = cconj(res[r * M + ((M - k) % M)]);
        cpx<R> E = {(R)0.5 * (Z.x + Zc.x), (R)0.5 * (Z.y + Zc.y)};
        cpx<R> D = {(R)0.5 * (Z.x - Zc.x), (R)0.5 * (Z.y - Zc.y)};
        // X = E - i W^k D
        cpx<R> WD = cmul(twL[k % Lx], D);
        cpx<R> X = {E.x + WD.y, E.y - WD.x};
        S[(size_t)(y0 + r) * H + k] = X;
    }
}

// column stage: forward column FFT of each input, spectral product (mode 0), spectrum store
// (mode 1) or reciprocal (mode 2), inverse column FFT -> T[c][y*H + kx]
template <typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_colmul(const ConvItem* items, const int* gstart,
                                                        int nitems)
{
    extern __shared__ char lds_raw[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const ConvItem& I = items[it];
    const int Lx = I.Lx, M = Lx / 2, H = M + 1, Ly = I.Ly;
    const int nrow = I.mode == 0 ? I.my : I.Ly;
    const int nrow_out = I.mode == 0 ? I.my : I.Ly;
    const int nin = I.nin, nout = I.mode == 0 ? I.nout : nin;
    const int nc = nin > nout ? nin : nout;
    const int cpg = conv_cols(I, nc);
    int kx0 = (blockIdx.x - gstart[it]) * cpg;
    int ncol = min(cpg, H - kx0);
    cpx<R>* a = (cpx<R>*)lds_raw;
    cpx<R>* b = a + (size_t)nc * cpg * Ly;
    const cpx<R>* S = (const cpx<R>*)I.wsS;
    for (int idx = threadIdx.x; idx < nin * ncol * Ly; idx += blockDim.x) {
        int tr = idx / Ly, y = idx - tr * Ly;
        int j = tr / ncol, kk = tr - j * ncol;
        a[idx] = y < nrow ? S[((size_t)j * nrow + y) * H + kx0 + kk] : cpx<R>{(R)0, (R)0};
    }
    __syncthreads();
    FftPlan plan = make_plan(Ly);
    const cpx<R>* twy = (const cpx<R>*)I.twLy;
    cpx<R>* f = fft_run<R>(a, b, plan, nin * ncol, (R)-1, twy, threadIdx.x, blockDim.x, LdsSync<R>());
    cpx<R>* o = f == a ? b : a;
    if (I.mode == 1) {
        // spectrum store: spec[c] layout [kx*Ly + ky], scaled
        for (int idx = threadIdx.x; idx < nin * ncol * Ly; idx += blockDim.x) {
            int tr = idx / Ly, ky = idx - tr * Ly;
            int j = tr / ncol, kk = tr - j * ncol;
            cpx<R> v = f[idx];
            v.x *= (R)I.scale;
            v.y *= (R)I.scale;
            ((cpx<R>*)I.out[j])[(size_t)(kx0 + kk) * Ly + ky] = v;
        }
        return;
    }
    for (int idx = threadIdx.x; idx < ncol * Ly; idx += blockDim.x) {
        int kk = idx / Ly, ky = idx - kk * Ly;
        int kx = kx0 + kk;
        cpx<R> xin[3];
        for (int j = 0; j < nin; j++) xin[j] = f[(j * ncol + kk) * Ly + ky];
        if (I.mode == 2) {
            for (int j = 0; j < nin; j++) {
                double re = (double)xin[j].x, im = (double)xin[j].y;
                double d = re * re + im * im;
                o[(j * ncol + kk) * Ly + ky] = {(R)(re / d), (R)(-im / d)};
            }
        } else {
            for (int i = 0; i < nout; i++) {
                cpx<R> acc = {(R)0, (R)0};
                for (int j = 0; j < nin; j++) {
                    const cpx<R>* sp = (const cpx<R>*)I.spec[i][j];
                    if (!sp) continue;
                    acc = cadd(acc, cmul(sp[(size_t)kx * Ly + ky], xin[j]));
                }
                o[(i * ncol + kk) * Ly + ky] = acc;
            }
        }
    }
    __syncthreads();
    cpx<R>* g = fft_run<R>(o, f, plan, nout * ncol, (R)1, twy, threadIdx.x, blockDim.x, LdsSync<R>());
    cpx<R>* T = (cpx<R>*)I.wsT;
    for (int idx = threadIdx.x; idx < nout * ncol * nrow_out; idx += blockDim.x) {
        int tr = idx / nrow_out, y = idx - tr * nrow_out;
        int i = tr / ncol, kk = tr - i * ncol;
        T[((size_t)i * nrow_out + y) * H + kx0 + kk] = g[(i * ncol + kk) * Ly + y];
    }
}

// inverse real row transforms T[c] -> out[c] rows, masked write (mode 0) or dense (mode 2)
template <typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_rowinv(const ConvItem* items, const int* gstart,
                                                        int nitems)
{
    extern __shared__ char lds_raw[];
    int it = find_item(gstart, nitems, blockIdx.x);
    const ConvItem& I = items[it];
    const int Lx = I.Lx, M = Lx / 2, H = M + 1;
    const int nrow = I.mode == 0 ? I.my : I.Ly;
    const int nxr = I.mode == 0 ? I.mx : I.Lx;
    const int rpg = conv_rows(I);
    const int ngy = (nrow + rpg - 1) / rpg;
    int t = blockIdx.x - gstart[it];
    int c = t / ngy;
    int y0 = (t - c * ngy) * rpg;
    int nr = min(rpg, nrow - y0);
    cpx<R>* a = (cpx<R>*)lds_raw;
    cpx<R>* b = a + rpg * M;
    const cpx<R>* T = (const cpx<R>*)I.wsT + (size_t)c * nrow * H;
    const cpx<R>* twL = (const cpx<R>*)I.twLx;
    for (int idx = threadIdx.x; idx < nr * M; idx += blockDim.x) {
        int r = idx / M, k = idx - r * M;
        const cpx<R>* row = T + (size_t)(y0 + r) * H;
        cpx<R> X = row[k];
        cpx<R> Xc = cconj(row[M - k]);
        cpx<R> E = {(R)0.5 * (X.x + Xc.x), (R)0.5 * (X.y + Xc.y)};
        cpx<R> D = {(R)0.5 * (X.x - Xc.x), (R)0.5 * (X.y - Xc.y)};
        cpx<R> O = cmul(D, cconj(twL[k]));
        a[idx] = {E.x - O.y, E.y + O.x};   // E + i O
    }
    __syncthreads();
    FftPlan plan = make_plan(M);
    cpx<R>* z = fft_run<R>(a, b, plan, nr, (R)1, (const cpx<R>*)I.twM, threadIdx.x, blockDim.x,
                           LdsSync<R>());
    double* dst = I.out[c];
    const double* add = I.addend[c];
    for (int idx = threadIdx.x; idx < nr * nxr; idx += blockDim.x) {
        int r = idx / nxr, x = idx - r * nxr;
        size_t ii = (size_t)(y0 + r) * nxr + x;
        if (I.mode == 0 && I.mask != M_ALL && !in_mask(I.mask, I.el[ii])) continue;
        cpx<R> zz = z[r * M + (x >> 1)];
        double u = (double)((x & 1) ? zz.y : zz.x);
        if (I.mode == 2) u *= I.scale;
        if (add) u += add[ii];
        dst[ii] = u;
    }
}

// ------------------------------------------------------------------------------------------
// influence coefficients (elascf_pcwcns, src/m_visc.f90:431-604) and their spectra
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(ELEM_THREADS) k_elascf(const CoefItem* items, const int* gstart, int nitems)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const CoefItem& I = items[it];
    const int mx = I.mx, my = I.my;
    const double PI = 3.14159265358979323846;
    const double e1 = (1.0 - I.nuv) / PI, e2 = 1.0 / PI, e3 = I.nuv / PI;
    const double tolx = I.dx * 1e-10, tolx2 = I.dx * 1e-20;
    const size_t nb = (size_t)(2 * mx) * (2 * my);
    int iy0 = my == 1 ? -my + 1 : -my;
    size_t k = (size_t)(blockIdx.x - gstart[it]) * blockDim.x + threadIdx.x;
    if (k < nb) {
        int ix = (int)(k % (2 * mx)) - mx, iy = (int)(k / (2 * mx)) - my;
        double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (iy >= iy0) {
            auto X = [&](int jx) { return (double)(float)jx * I.dx + I.xshft - 0.5 * I.dx; };
            auto Y = [&](int jy) { return (double)(float)jy * I.dy + I.yshft - 0.5 * I.dy; };
            auto Rr = [&](int jx, int jy) { double x = X(jx), y = Y(jy); return sqrt(x * x + y * y); };
            // cf21 = nu/pi * (-r00 + r01 + r10 - r11)
            double c21 = 0.0;
            c21 = c21 - e3 * Rr(ix, iy);
            c21 = c21 + e3 * Rr(ix, iy + 1);
            c21 = c21 + e3 * Rr(ix + 1, iy);
            c21 = c21 - e3 * Rr(ix + 1, iy + 1);
            auto xly = [&](int jx, int jy) {   // x log((y1+r1)/(y0+r0)) over [jy, jy+1]
                double x = X(jx), y0 = Y(jy), y1 = Y(jy + 1);
                double r0 = sqrt(x * x + y0 * y0), r1 = sqrt(x * x + y1 * y1);
                if (fmin(fabs(y0 + r0), fabs(y1 + r1)) < tolx) return 0.0;
                return x * log((y1 + r1) / (y0 + r0));
            };
            auto ylx = [&](int jx, int jy) {
                double y = Y(jy), x0 = X(jx), x1 = X(jx + 1);
                double r0 = sqrt(x0 * x0 + y * y), r1 = sqrt(x1 * x1 + y * y);
                if (fmin(fabs(x0 + r0), fabs(x1 + r1)) < tolx) return 0.0;
                return y * log((x1 + r1) / (x0 + r0));
            };
            double xa = xly(ix, iy), xb = xly(ix + 1, iy), ya = ylx(ix, iy), yb = ylx(ix, iy + 1);
            double c33 = 0.0, c11 = 0.0, c22 = 0.0;
            c33 = c33 - e1 * (xa + ya);
            c33 = c33 + e1 * (xb + yb);
            c11 = c11 - (e1 * xa + e2 * ya);
            c11 = c11 + (e1 * xb + e2 * yb);
            c22 = c22 - (e1 * ya + e2 * xa);
            c22 = c22 + (e1 * yb + e2 * xb);
            double c13 = 0.0, c23 = 0.0;
            if (fabs(I.akv) >= 1e-6) {
                auto t13 = [&](int jx, int jy) {
                    double x = X(jx), y = Y(jy), r = sqrt(x * x + y * y);
                    return y * log(r) + x * atan(y / (x + tolx2));
                };
                auto t23 = [&](int jx, int jy) {
                    double x = X(jx), y = Y(jy), r = sqrt(x * x + y * y);
                    return x * log(r) + y * atan(x / (y + tolx2));
                };
                double ek = e2 * I.akv;
                c13 = c13 - ek * t13(ix, iy);
                c13 = c13 + ek * t13(ix, iy + 1);
                c13 = c13 + ek * t13(ix + 1, iy);
                c13 = c13 - ek * t13(ix + 1, iy + 1);
                c23 = c23 - ek * t23(ix, iy);
                c23 = c23 + ek * t23(ix, iy + 1);
                c23 = c23 + ek * t23(ix + 1, iy);
                c23 = c23 - ek * t23(ix + 1, iy + 1);
            }
            // the xy and yx blocks hold one value (k_stdygs3 stages only the xy block)
            c[0] = c11; c[1] = c21; c[2] = c13;
            c[3] = c21; c[4] = c22; c[5] = c23;
            c[6] = -c13; c[7] = -c23; c[8] = c33;
        }
        for (int b = 0; b < 9; b++) I.cf[(size_t)b * nb + k] = c[b];
    }
}

// rolling: cv keeps the normal row of cs; csv = cs - cv, cv = cs - csv on the tangential rows
// (src/m_visc.f90:319-358 with facdq = 1)
__global__ void k_rollcoef(int n, const double* cs, double* cv)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < (size_t)n; k += (size_t)gridDim.x * blockDim.x) {
        for (int jk = 0; jk < 3; jk++) {
            cv[(size_t)(6 + jk) * n + k] = cs[(size_t)(6 + jk) * n + k];
            for (int ik = 0; ik < 2; ik++) {
                size_t b = (size_t)(ik * 3 + jk) * n + k;
                double csv = cs[b] - cv[b];
                cv[b] = cs[b] - csv;
            }
        }
    }
}

__global__ void __launch_bounds__(ELEM_THREADS) k_embed(const EmbedItem* items, const int* gstart, int nitems)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const EmbedItem& I = items[it];
    size_t n = (size_t)I.Lx * I.Ly;
    size_t k = (size_t)(blockIdx.x - gstart[it]) * blockDim.x + threadIdx.x;
    if (k < n) {
        int x = (int)(k % I.Lx), y = (int)(k / I.Lx);
        int dx = x < I.Lx / 2 ? x : x - I.Lx;
        int dy = y < I.Ly / 2 ? y : y - I.Ly;
        double v = 0.0;
        if (dx >= -I.rx && dx < I.rx && dy >= -I.ry && dy < I.ry && dx >= -I.cf_mx && dx < I.cf_mx &&
            dy >= -I.cf_my && dy < I.cf_my)
            v = I.scale * I.cf[(size_t)(dy + I.cf_my) * (2 * I.cf_mx) + (dx + I.cf_mx)];
        I.out[k] = v;
    }
}

__global__ void __launch_bounds__(ELEM_THREADS) k_extract(const ExtractItem* items, const int* gstart, int nitems)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const ExtractItem& I = items[it];
    size_t n = (size_t)(2 * I.cf_mx) * (2 * I.cf_my);
    size_t k = (size_t)(blockIdx.x - gstart[it]) * blockDim.x + threadIdx.x;
    if (k < n) {
        int dx = (int)(k % (2 * I.cf_mx)) - I.cf_mx, dy = (int)(k / (2 * I.cf_mx)) - I.cf_my;
        int x = (dx + I.Lx) % I.Lx, y = (dy + I.Ly) % I.Ly;
        I.cf[k] = I.scale * I.in[(size_t)y * I.Lx + x];
    }
}

__global__ void k_to_float(size_t n, const double* in, float* out)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        out[k] = (float)in[k];
}

// ------------------------------------------------------------------------------------------
// per-case fields and force sums
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(ELEM_THREADS) k_setup(const SetupItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const SetupItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[2] = {-INFINITY, -INFINITY};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int ix = i % I.mx, iy = i / I.mx;
        double x = I.xc1 + ix * I.dx, y = I.yc1 + iy * I.dy;
        // previous-time tractions (P digit)
        if (I.prev) {
            I.pvn[i] = I.pn[i];
            I.pvx[i] = I.px[i];
            I.pvy[i] = I.py[i];
        } else {
            I.pvn[i] = 0.0;
            I.pvx[i] = 0.0;
            I.pvy[i] = 0.0;
        }
        double h;
        if (I.ibase == 1) {
            h = I.prm[0] * x * x + I.prm[1] * x * y + I.prm[2] * y * y + I.prm[3] * x + I.prm[4] * y + I.prm[5];
            I.h[i] = h;
        } else {
            h = I.h[i];
        }
        // rigid slip w = (xi - phi (y + yofs), eta + phi (x + xofs)) + exrhs, times -dq
        // (set_tang_rhs, src/m_sdis.f90:556-579)
        const double ex = I.exrhs ? I.hx[i] : 0.0, ey = I.exrhs ? I.hy[i] : 0.0;
        double vx = I.facx * -(y + I.yofs - 0.0) * I.cphi + ex;
        double vy = I.facy * (x + I.xofs - 0.0) * I.cphi + ey;
        vx += I.cksi;
        vy += I.ceta;
        I.hx[i] = vx * -I.dq;
        I.hy[i] = vy * -I.dq;
        // initial estimate
        if (I.zero_pn) I.pn[i] = 0.0;
        if (I.zero_pt) {
            I.px[i] = 0.0;
            I.py[i] = 0.0;
        }
        if (I.hertz) {
            double ax = x / I.hz_aa, by = y / I.hz_bb;
            double f = I.hertz == 1 ? 1.0 - ax * ax - by * by : 1.0 - ax * ax;
            if (I.hertz == 2 && fabs(y) > I.hz_bb) f = 0.0;
            double pn = I.pnmax * sqrt(fmax(0.0, f));
            I.pn[i] = pn;
            I.el[i] = pn > (double)1e-20f ? ADHES : EXTER;
        }
        if (I.regularize) {
            // I=1 (src/m_sdis.f90:722-735): exterior tractions 0, slip tractions on the bound
            const int8_t e = I.el[i];
            if (e <= EXTER) {
                I.pn[i] = 0.0;
                I.px[i] = 0.0;
                I.py[i] = 0.0;
            } else if (e == SLIP) {
                const double pabs = fmax(1e-10, sqrt(I.px[i] * I.px[i] + I.py[i] * I.py[i]));
                I.px[i] = I.mu * I.pn[i] * I.px[i] / pabs;
                I.py[i] = I.mu * I.pn[i] * I.py[i] / pabs;
            }
        }
        if (h > (double)1e29f) I.el[i] = EXTER;
        acc[0] = fmax(acc[0], h);
        acc[1] = fmax(acc[1], -h);
    }
    write_partials<2>(acc, 1, ro, gstart, it);
}

__global__ void __launch_bounds__(ELEM_THREADS) k_forces(const SetupItem* items, const int* gstart,
                                                         int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const SetupItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        int ix = i % I.mx, iy = i / I.mx;
        double x = I.xc1 + ix * I.dx, y = I.yc1 + iy * I.dy;
        acc[0] += I.pn[i];
        acc[1] += I.px[i];
        acc[2] += I.py[i];
        acc[3] += I.px[i] * y;
        acc[4] += I.py[i] * x;
    }
    write_partials<5>(acc, 0, ro, gstart, it);
}


// ------------------------------------------------------------------------------------------
// sensitivity solves: state copies, perturbed right-hand sides, contact extents
// (src/m_solvpn.f90:531-650, src/m_solvpt.f90:497-796)
// ------------------------------------------------------------------------------------------

__global__ void __launch_bounds__(ELEM_THREADS) k_sens(const SensItem* items, const int* gstart,
                                                       int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const SensItem& I = items[it];
    int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    double acc[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        if (I.mode == 0) {
            if (I.el_dst) I.el_dst[i] = I.el_src[i];
            for (int v = 0; v < I.nvec; v++) I.dst[v][i] = I.src[v][i];
        } else if (I.mode == 1) {
            // wsloc = wstot + facdt dcrp dq (facdt = 1 on the interior), src/m_solvpt.f90:644-690
            double wx = I.wbx[i], wy = I.wby[i];
            if (I.el[i] >= ADHES) {
                int ix = i % I.mx, iy = i / I.mx;
                double x = I.xc1 + ix * I.dx, y = I.yc1 + iy * I.dy;
                // facdt < 1 only near the leading edge of transient rolling with dq > dx
                const double d = I.facdt ? I.facdt[i] * I.d : I.d;
                if (I.which == 0) wx = wx + d * I.dq;
                else if (I.which == 1) wy = wy + d * I.dq;
                else {
                    wx = wx - d * y * I.dq;
                    wy = wy + d * x * I.dq;
                }
            }
            I.wsx[i] = wx;
            I.wsy[i] = wy;
        } else if (I.el[i] >= ADHES) {
            double ix = (double)(i % I.mx + 1), iy = (double)(i / I.mx + 1);
            acc[0] = fmax(acc[0], -ix);
            acc[1] = fmax(acc[1], ix);
            acc[2] = fmax(acc[2], -iy);
            acc[3] = fmax(acc[3], iy);
        }
    }
    if (I.mode == 2) write_partials<4>(acc, 1, ro, gstart, it);
}

// leading edge of transient rolling, sxbnd and stang_rhs (1b) (src/m_leadedge.f90:92-394,
// src/m_stang.f90:838-935); see LedgeItem
__global__ void __launch_bounds__(ELEM_THREADS) k_ledge(const LedgeItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const LedgeItem& I = items[it];
    (void)ro;
    if (I.mode == 0) {
        const int iy = (blockIdx.x - gstart[it]) * ELEM_THREADS + threadIdx.x;
        if (iy >= I.my) return;
        const int mx = I.mx;
        const size_t r0 = (size_t)iy * mx;
        int ix0 = -1;   // nearest leading-edge position at or after ix (interior, exterior next)
        for (int ix = mx - 1; ix >= 0; ix--) {
            const size_t ii = r0 + ix;
            const int8_t e = I.el[ii];
            if (e >= ADHES && (ix == mx - 1 || I.el[ii + 1] <= EXTER)) ix0 = ix;
            double f = 1.0, l = -1.0;
            if (e <= EXTER) {
                f = 0.0;
            } else if (ix + 3 <= mx) {   // 1-based ix + 2 <= mx: room to estimate the edge
                const double xbnd = (I.xc1 + ix0 * I.dx) + I.fxdfac * I.dx;
                f = fmin(1.0, (xbnd - (I.xc1 + ix * I.dx)) / I.dq);
                if (f < 0.9999) l = (double)ix0;
            }
            I.facdt[ii] = f;
            I.lix[ii] = l;
        }
        return;
    }
    const int base = (blockIdx.x - gstart[it]) * ELEM_THREADS * ELEM_PER_THREAD;
    for (int e = 0; e < ELEM_PER_THREAD; e++) {
        const int i = base + e * ELEM_THREADS + threadIdx.x;
        if (i >= I.n) break;
        double wx = 0.0, wy = 0.0;
        if (I.el[i] >= ADHES && I.steady) {
            // steady rolling (stang_rhs (1a), (1b) with T=3): the rigid slip times facdt; the
            // displacement at the leading edge enters the solver's shift instead
            wx = -I.facdt[i] * I.hx[i];
            wy = -I.facdt[i] * I.hy[i];
        } else if (I.el[i] >= ADHES) {
            const double f = I.facdt[i];
            wx = -f * I.hx[i];
            wy = -f * I.hy[i];
            const int j0 = (int)I.lix[i];
            if (j0 < 0) {
                wx = wx - I.ux[i];
                wy = wy - I.uy[i];
            } else if (j0 + 3 <= I.mx) {   // subnd: 1-based ix0 + 2 <= mx, else ubnd = 0
                const size_t r = (size_t)(i / I.mx) * I.mx + j0;
                wx = wx - ((1.0 - I.fxdfac) * I.bx[r] + I.fxdfac * I.bx[r + 1]);
                wy = wy - ((1.0 - I.fxdfac) * I.by[r] + I.fxdfac * I.by[r + 1]);
            }
        }
        I.wsx[i] = wx;
        I.wsy[i] = wy;
    }
}

// ------------------------------------------------------------------------------------------
// SteadyGS, src/m_solvpt.f90:2825-3254 (chi = 0, no plasticity, no third body)
// ------------------------------------------------------------------------------------------

#ifndef GS_THREADS_N
#define GS_THREADS_N 512
#endif
constexpr int GS_THREADS = GS_THREADS_N;

// Rows with an interior element in the trailing-edge column ix = 1: steady rolling then
// needs ConvexGS (src/m_stang.f90:131-143, 185-190).  One thread per row.  Also returns the
// tangential coefficients of element (0,0) (blocks xx, xy, yx, yy, unscaled).
__global__ void __launch_bounds__(ELEM_THREADS) k_guard(const GsItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsItem& I = items[it];
    int iy = (blockIdx.x - gstart[it]) * ELEM_THREADS + threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (iy < I.my && I.el[(size_t)iy * I.mx] >= ADHES) acc[0] = 1.0;
    if (iy == 0) {
        const size_t k0 = (size_t)I.my * 2 * I.mx + I.mx;
        acc[1] = I.cf[0][0][k0];
        acc[2] = I.cf[0][1][k0];
        acc[3] = I.cf[1][0][k0];
        acc[4] = I.cf[1][1][k0];
    }
    write_partials<5>(acc, 0, ro, gstart, it);
}

// One work-group per patch runs all sweeps.  Per interior element (rows in order, each row
// from the trailing edge ix = 1 to the leading edge): the slip due to all differences dp by
// direct summation over the row extents of the contact area (gf3_AijPj, all threads), the
// local 2x2 problem with the coefficients relative to the first upstream non-adhesion
// element (thread 0), then the upstream part of the row is rebuilt from dp with the traction
// bound applied (:3086-3127).  The rebuild stops at the first upstream element that is not
// in adhesion: the elements beyond it depend on its unchanged traction only.  The current
// row lives in LDS; dp is written through to global memory for the summation.
// Returns (itgs, dif, algorithmic flops: 8 per coefficient-difference product of the sums,
// 0, 0, dif of the first sweep, 0, 0, 0, 0).
__global__ void __launch_bounds__(GS_THREADS) k_stdygs(const GsItem* items, const int* gstart,
                                                       int nitems, RedOut ro)
{
    extern __shared__ double gs_lds[];
    __shared__ double red[16 * MAXK];
    __shared__ double s_ctl[4];
    __shared__ int s_len;
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsItem& I = items[it];
    const int mx = I.mx, my = I.my, n = I.n, tid = threadIdx.x, T = blockDim.x;
    double* rpx = gs_lds;
    double* rpy = rpx + mx;
    double* rdpx = rpy + mx;
    double* rdpy = rdpx + mx;
    double* rg = rdpy + mx;
    int* roff = (int*)(rg + mx);          // my + 1 row offsets into the summation list
    int* rj0 = roff + my + 1;             // first jx (0-based) of each row's range
    int8_t* rel = (int8_t*)(rj0 + my);
    double* dpx = I.dpx;
    double* dpy = I.dpy;
    const double* c00 = I.cf[0][0];
    const double* c01 = I.cf[0][1];
    const double* c10 = I.cf[1][0];
    const double* c11 = I.cf[1][1];

    // dp = backward differences along x; the leading-edge column keeps p (:2893-2908)
    for (int i = tid; i < n; i += T) {
        int ix = i % mx;
        if (ix != mx - 1) {
            dpx[i] = I.px[i] - I.px[i + 1];
            dpy[i] = I.py[i] - I.py[i + 1];
        } else {
            dpx[i] = I.px[i];
            dpy[i] = I.py[i];
        }
    }
    // summation ranges jx in [row1st-1, rowlst+1] per row (src/m_aijpj.f90:173-185)
    for (int jy = tid; jy < my; jy += T) {
        const int8_t* row = I.el + (size_t)jy * mx;
        int first = 0;
        while (first < mx - 1 && row[first] <= EXTER) first++;
        int last = -1;
        for (int ix = first; ix < mx; ix++)
            if (row[ix] >= ADHES) last = ix;
        int j0 = first - 1 < 0 ? 0 : first - 1;
        int j1 = last + 1 > mx - 1 ? mx - 1 : last + 1;
        rj0[jy] = j0;
        roff[jy + 1] = j1 >= j0 ? j1 - j0 + 1 : 0;
    }
    __syncthreads();
    if (tid == 0) {
        roff[0] = 0;
        for (int jy = 0; jy < my; jy++) roff[jy + 1] += roff[jy];
        s_len = roff[my];
    }
    __syncthreads();
    const int len = s_len;
    for (int jy = tid; jy < my; jy += T) {
        int j0 = rj0[jy], cnt = roff[jy + 1] - roff[jy];
        for (int q = 0; q < cnt; q++) {
            int jx = j0 + q;
            I.list[2 * (roff[jy] + q)] = jy * mx + jx;
            I.list[2 * (roff[jy] + q) + 1] = jy * 2 * mx + jx;
        }
    }
    __syncthreads();

    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0;
    while (dif >= difid && itgs < I.maxgs) {
        itgs++;
        double difacc = 0.0;   // thread 0
        for (int iy = 0; iy < my; iy++) {
            const size_t r0 = (size_t)iy * mx;
            for (int ix = tid; ix < mx; ix += T) {
                rpx[ix] = I.px[r0 + ix];
                rpy[ix] = I.py[r0 + ix];
                rdpx[ix] = dpx[r0 + ix];
                rdpy[ix] = dpy[r0 + ix];
                rg[ix] = I.mu * I.pn[r0 + ix];
                rel[ix] = I.el[r0 + ix];
            }
            __syncthreads();
            for (int ix = 0; ix < mx; ix++) {
                if (rel[ix] < ADHES) continue;
                // s = ws + A dp at (ix, iy)
                const int base = (iy + my) * 2 * mx + ix + mx;
                double acc[2] = {0.0, 0.0};
                for (int e = tid; e < len; e += T) {
                    int j = I.list[2 * e], c = base - I.list[2 * e + 1];
                    double qx = dpx[j], qy = dpy[j];
                    acc[0] += c00[c] * qx + c01[c] * qy;
                    acc[1] += c10[c] * qx + c11[c] * qy;
                }
                group_reduce<2>(acc, 0, red);
                if (tid == 0) {
                    const size_t ii = r0 + ix;
                    int jx = ix - 1;
                    while (jx > 0 && rel[jx] == ADHES) jx--;
                    const int b0 = my * 2 * mx + mx, bj = b0 + (jx - ix);
                    double a11 = I.ga_inv * (c00[b0] - c00[bj]);
                    double a12 = I.ga_inv * (c01[b0] - c01[bj]);
                    double a21 = I.ga_inv * (c10[b0] - c10[bj]);
                    double a22 = I.ga_inv * (c11[b0] - c11[bj]);
                    double sv[2] = {I.wsx[ii] + I.ga_inv * acc[0], I.wsy[ii] + I.ga_inv * acc[1]};
                    double pr[3] = {rpx[ix], rpy[ix], I.pn[ii]};
                    int8_t e8 = rel[ix];
                    elmtrc(e8, a11, a12, a21, a22, I.eps, pr, I.mu, sv, I.omegah, I.omegas);
                    double ddx = pr[0] - rpx[ix], ddy = pr[1] - rpy[ix];
                    difacc += ddx * ddx + ddy * ddy;
                    rdpx[ix] += ddx;
                    rdpy[ix] += ddy;
                    dpx[ii] = rdpx[ix];
                    dpy[ii] = rdpy[ix];
                    rpx[ix] = pr[0];
                    rpy[ix] = pr[1];
                    rel[ix] = e8;
                    I.sx[ii] = sv[0];
                    I.sy[ii] = sv[1];
                    for (int kx = ix - 1; kx >= 0; kx--) {
                        int8_t ek = rel[kx];
                        if (ek == ADHES) {
                            double qx = rpx[kx + 1] + rdpx[kx], qy = rpy[kx + 1] + rdpy[kx];
                            double ptabs = sqrt(qx * qx + qy * qy), ptbnd = rg[kx];
                            if (ptabs > ptbnd) {
                                qx = qx * ptbnd / ptabs;
                                qy = qy * ptbnd / ptabs;
                                rdpx[kx] = qx - rpx[kx + 1];
                                rdpy[kx] = qy - rpy[kx + 1];
                                dpx[r0 + kx] = rdpx[kx];
                                dpy[r0 + kx] = rdpy[kx];
                            }
                            rpx[kx] = qx;
                            rpy[kx] = qy;
                        } else if (ek == SLIP || ek == PLAST) {
                            rdpx[kx] = rpx[kx] - rpx[kx + 1];
                            rdpy[kx] = rpy[kx] - rpy[kx + 1];
                            dpx[r0 + kx] = rdpx[kx];
                            dpy[r0 + kx] = rdpy[kx];
                            break;
                        } else {
                            if (rel[kx + 1] >= ADHES) {
                                rdpx[kx] = -rpx[kx + 1];
                                rdpy[kx] = -rpy[kx + 1];
                                dpx[r0 + kx] = rdpx[kx];
                                dpy[r0 + kx] = rdpy[kx];
                            }
                            break;
                        }
                    }
                }
                __syncthreads();
            }
            for (int ix = tid; ix < mx; ix += T) {
                I.px[r0 + ix] = rpx[ix];
                I.py[r0 + ix] = rpy[ix];
                I.el[r0 + ix] = rel[ix];
            }
            __syncthreads();
        }
        // dif = sqrt(sum |dp_k|^2 / 2k); difid from the rms of p over all elements
        double sq[1] = {0.0};
        if (I.p0x) {
            // sensitivity solve: rms of the change from the initial estimate (:3138-3144)
            for (int i = tid; i < n; i += T) {
                double ax = I.px[i] + -1.0 * I.p0x[i], ay = I.py[i] + -1.0 * I.p0y[i];
                sq[0] += ax * ax + ay * ay;
            }
        } else {
            for (int i = tid; i < n; i += T) sq[0] += I.px[i] * I.px[i] + I.py[i] * I.py[i];
        }
        group_reduce<1>(sq, 0, red);
        if (tid == 0) {
            double rms = sqrt(sq[0] / (double)(2 * n > 1 ? 2 * n : 1));
            s_ctl[0] = sqrt(difacc / (2.0 * I.k));
            s_ctl[1] = I.eps * fmax(1e-6, I.facnel * rms);
        }
        __syncthreads();
        dif = s_ctl[0];
        difid = s_ctl[1];
        if (itgs == 1) dif1 = dif;
        __syncthreads();
    }
    if (tid == 0) {
        ro.partials[(size_t)blockIdx.x * 10] = (double)itgs;
        ro.partials[(size_t)blockIdx.x * 10 + 1] = dif;
        ro.partials[(size_t)blockIdx.x * 10 + 2] = 8.0 * itgs * (double)I.k * (double)len;
        for (int q = 3; q < 10; q++) ro.partials[(size_t)blockIdx.x * 10 + q] = 0.0;
        ro.partials[(size_t)blockIdx.x * 10 + 5] = dif1;
    }
}


// SteadyGS for rows of up to ~1000 elements: the same sweeps as k_stdygs, with the influence
// of the other rows on the current row summed once at the start of the row by all threads
// (each source row's coefficient row and differences staged through LDS, alternating between
// two buffers, one barrier per source row), so that the sequential part along the row only adds the
// current row's own influence and runs on one wavefront without work-group barriers.
// The row sums are Toeplitz products along x: a thread owns GS_TB consecutive targets and
// takes the sources in chunks of GS_QB, so that one window of GS_TB + GS_QB - 1 coefficients
// per block serves GS_TB x GS_QB products from registers.
// Summation order differs from gf3_AijPj; the sums are the same.  Returns (itgs, dif,
// algorithmic flops, wall-clock ticks (100 MHz) spent in the row sums and in the sequential
// parts, dif of the first sweep, and -- built with GS_PHASE_TIMING -- the ticks of the four
// steps of the sequential part).

// LDS of k_stdygs2 without the staging area; the staging area holds `sr` source rows
__host__ __device__ inline size_t gs2_base_lds(int mx, int my)
{
    return (size_t)(10 * mx) * sizeof(double) + 2 * GS_THREADS * sizeof(double) +
           (size_t)4 * my * sizeof(int) + mx + 64;
}
#ifndef GS_TB_N
#define GS_TB_N 8
#endif
#ifndef GS_QB_N
#define GS_QB_N 4
#endif
constexpr int GS_TB = GS_TB_N, GS_QB = GS_QB_N;
// doubles of one staging buffer: 4 coefficient windows of up to 2 mx + GS_TB + GS_QB values
// (each padded to a multiple of 4) and 2 x (mx + GS_QB) differences; this also holds the
// 4 x 2 mx values of the own-row coefficient row
__host__ __device__ inline int gs2_window(int mx) { return (2 * mx + GS_TB + GS_QB + 3) / 4 * 4; }
__host__ __device__ inline int gs2_stage_stride(int mx)
{
    return (4 * gs2_window(mx) + 2 * (mx + GS_QB) + 7) / 8 * 8;
}
__host__ __device__ inline size_t gs2_row_lds(int mx) { return (size_t)gs2_stage_stride(mx) * sizeof(double); }
// two staging buffers: the next source row is written while others still read the current
// source rows staged per barrier (GS_R, when two groups of them fit next to the fixed part)
constexpr int GS_R = 1;
__host__ __device__ inline int gs2_rows_per_stage(int mx, int my)
{
    return gs2_base_lds(mx, my) + (size_t)2 * GS_R * gs2_row_lds(mx) <= MAX_LDS ? GS_R : 1;
}
__host__ __device__ inline int gs2_stage_rows(int mx, int my) { return 2 * gs2_rows_per_stage(mx, my); }
__host__ __device__ inline size_t gs2_lds(int mx, int my)
{
    return gs2_base_lds(mx, my) + (size_t)gs2_stage_rows(mx, my) * gs2_row_lds(mx);
}

__global__ void __launch_bounds__(GS_THREADS) k_stdygs2(const GsItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    extern __shared__ double gs_lds[];
    __shared__ double red_s[16 * MAXK];
    __shared__ double s_ctl[4];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsItem& I = items[it];
    const int mx = I.mx, my = I.my, n = I.n, tid = threadIdx.x, T = blockDim.x;
    const int M2 = 2 * mx;
    double* rpx = gs_lds;
    double* rpy = rpx + mx;
    double* rdpx = rpy + mx;
    double* rdpy = rdpx + mx;
    double* rg = rdpy + mx;
    double* rpn = rg + mx;
    double* rwx = rpn + mx;
    double* rwy = rwx + mx;
    double* ucx = rwy + mx;                // sums over the other rows, per column
    double* ucy = ucx + mx;
    double* red = ucy + mx;                // 2 partial sums per thread
    int* rj0 = (int*)(red + 2 * GS_THREADS);   // summation range per row
    int* rj1 = rj0 + my;
    int* ri0 = rj1 + my;                   // interior range per row
    int* ri1 = ri0 + my;
    int8_t* rel = (int8_t*)(ri1 + my);
    // staging area after the fixed part, 64-byte aligned (an offset from the LDS base keeps the
    // compiler on LDS instructions)
    double* stg = gs_lds + ((10 * mx + 2 * GS_THREADS) * 8 + 4 * my * 4 + mx + 63) / 64 * 8;
    double* dpx = I.dpx;
    double* dpy = I.dpy;
    const double* cfb[4] = {I.cf[0][0], I.cf[0][1], I.cf[1][0], I.cf[1][1]};

    for (int i = tid; i < n; i += T) {
        int ix = i % mx;
        if (ix != mx - 1) {
            dpx[i] = I.px[i] - I.px[i + 1];
            dpy[i] = I.py[i] - I.py[i + 1];
        } else {
            dpx[i] = I.px[i];
            dpy[i] = I.py[i];
        }
    }
    for (int jy = tid; jy < my; jy += T) {
        const int8_t* row = I.el + (size_t)jy * mx;
        int first = 0;
        while (first < mx - 1 && row[first] <= EXTER) first++;
        int last = -1, lo = mx;
        for (int ix = 0; ix < mx; ix++)
            if (row[ix] >= ADHES) {
                last = ix;
                if (ix < lo) lo = ix;
            }
        rj0[jy] = first - 1 < 0 ? 0 : first - 1;
        rj1[jy] = last + 1 > mx - 1 ? mx - 1 : last + 1;
        ri0[jy] = lo;
        ri1[jy] = last;
    }
    __syncthreads();

    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0;
    unsigned long long t_rows = 0, t_inrow = 0;   // thread 0: wall-clock ticks of the two phases
    // with GS_PHASE_TIMING: ticks of the in-row steps (own-row sum, upstream search, element
    // problem, chain rebuild), thread 0
    unsigned long long ph[4] = {0, 0, 0, 0};
    while (dif >= difid && itgs < I.maxgs) {
        itgs++;
        double difacc = 0.0;   // thread 0
        for (int iy = 0; iy < my; iy++) {
            if (ri1[iy] < ri0[iy]) continue;
            const unsigned long long tr0 = wall_clock64();
            const int t0 = ri0[iy], nt = ri1[iy] - t0 + 1;
            const int ntg = (nt + GS_TB - 1) / GS_TB, ntp = ntg * GS_TB;
            const int P = T / ntg > 0 ? T / ntg : 1;
            const int tg = tid % ntg, part = tid / ntg;
            const bool active = part < P;
            const size_t r0 = (size_t)iy * mx;
            const int SW = gs2_stage_stride(mx);
            double ax[GS_TB], ay[GS_TB];
#pragma unroll
            for (int i = 0; i < GS_TB; i++) ax[i] = ay[i] = 0.0;
            // source rows staged R at a time into one of two groups of buffers used in turn: one
            // barrier per group suffices
            const int R = gs2_rows_per_stage(mx, my);
            int sgrp = 0, jy = 0;
            while (true) {
                int rows[GS_R], nr = 0;
                for (; jy < my && nr < R; jy++)
                    if (jy != iy && rj1[jy] >= rj0[jy]) rows[nr++] = jy;
                if (nr == 0) break;
                double* grp = stg + (size_t)sgrp * R * SW;
                sgrp ^= 1;
                for (int r = 0; r < nr; r++) {
                    // sources padded with zero differences to nch chunks; window S[k],
                    // k = t - q + Q - 1, holds the coefficient of target t0 + t and source j0 + q
                    // (zero outside the block)
                    const int sy = rows[r], j0 = rj0[sy], nq = rj1[sy] - j0 + 1;
                    const int Q = (nq + GS_QB - 1) / GS_QB * GS_QB;
                    const int Wn = ntp + Q - 1, Wp = (Wn + 1 + 3) / 4 * 4;
                    const int col0 = t0 - j0 - (Q - 1) + mx;
                    double* sc = grp + (size_t)r * SW;
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        const double* crow = cfb[bb] + (size_t)(iy - sy + my) * M2;
                        for (int k = tid; k < Wp; k += T) {
                            const int col = col0 + k;
                            sc[bb * Wp + k] = (k < Wn && col >= 0 && col < M2) ? crow[col] : 0.0;
                        }
                    }
                    double* sdx = sc + 4 * Wp;
                    double* sdy = sdx + Q;
                    for (int e = tid; e < Q; e += T) {
                        const bool in = e < nq;
                        sdx[e] = in ? dpx[(size_t)sy * mx + j0 + e] : 0.0;
                        sdy[e] = in ? dpy[(size_t)sy * mx + j0 + e] : 0.0;
                    }
                }
                __syncthreads();
                if (active) {
                    for (int r = 0; r < nr; r++) {
                        const int sy = rows[r], nq = rj1[sy] - rj0[sy] + 1;
                        const int nch = (nq + GS_QB - 1) / GS_QB, Q = nch * GS_QB;
                        const int Wp = (ntp + Q + 3) / 4 * 4;
                        const double* sc = grp + (size_t)r * SW;
                        const double* sdx = sc + 4 * Wp;
                        const double* sdy = sdx + Q;
                        for (int c = part; c < nch; c += P) {
                            const int kmin = tg * GS_TB - c * GS_QB + (Q - GS_QB);
                            double qx[GS_QB], qy[GS_QB];
#pragma unroll
                            for (int j = 0; j < GS_QB; j++) {
                                qx[j] = sdx[c * GS_QB + j];
                                qy[j] = sdy[c * GS_QB + j];
                            }
#pragma unroll
                            for (int h = 0; h < 2; h++) {
                                // blocks (0, 1) give the x sums, blocks (2, 3) the y sums
                                const double2* wa = (const double2*)(sc + (2 * h) * Wp + kmin);
                                const double2* wb = (const double2*)(sc + (2 * h + 1) * Wp + kmin);
                                double va[GS_TB + GS_QB], vb[GS_TB + GS_QB];
#pragma unroll
                                for (int m = 0; m < (GS_TB + GS_QB) / 2; m++) {
                                    const double2 u = wa[m], v = wb[m];
                                    va[2 * m] = u.x;
                                    va[2 * m + 1] = u.y;
                                    vb[2 * m] = v.x;
                                    vb[2 * m + 1] = v.y;
                                }
#pragma unroll
                                for (int i = 0; i < GS_TB; i++) {
                                    double acc = h == 0 ? ax[i] : ay[i];
#pragma unroll
                                    for (int j = 0; j < GS_QB; j++) {
                                        const int m = i + GS_QB - 1 - j;
                                        acc += va[m] * qx[j] + vb[m] * qy[j];
                                    }
                                    if (h == 0) ax[i] = acc;
                                    else ay[i] = acc;
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
            // sums over the parts, one target of each group per pass
            for (int i = 0; i < GS_TB; i++) {
                if (active) {
                    red[2 * (part * ntg + tg)] = ax[i];
                    red[2 * (part * ntg + tg) + 1] = ay[i];
                }
                __syncthreads();
                for (int g = tid; g < ntg; g += T) {
                    const int c = g * GS_TB + i;
                    if (c < nt) {
                        double sx = 0.0, sy = 0.0;
                        for (int p = 0; p < P; p++) {
                            sx += red[2 * (p * ntg + g)];
                            sy += red[2 * (p * ntg + g) + 1];
                        }
                        ucx[t0 + c] = sx;
                        ucy[t0 + c] = sy;
                    }
                }
                __syncthreads();
            }
            for (int ix = tid; ix < mx; ix += T) {
                rpx[ix] = I.px[r0 + ix];
                rpy[ix] = I.py[r0 + ix];
                rdpx[ix] = dpx[r0 + ix];
                rdpy[ix] = dpy[r0 + ix];
                double pn = I.pn[r0 + ix];
                rpn[ix] = pn;
                rg[ix] = I.mu * pn;
                rwx[ix] = I.wsx[r0 + ix];
                rwy[ix] = I.wsy[r0 + ix];
                rel[ix] = I.el[r0 + ix];
            }
            double* stc = stg;                     // coefficient row 0, offsets (q - mx, 0)
            for (int e = tid; e < 4 * M2; e += T) {
                int b = e / M2, q = e - b * M2;
                stc[e] = cfb[b][(size_t)my * M2 + q];
            }
            __syncthreads();
            const unsigned long long tr1 = wall_clock64();
            if (tid < 64) {
                const int lane = tid, j0 = rj0[iy], j1 = rj1[iy];
                // first upstream element not in adhesion (:2972-2977): the largest jx in [1, ix-1]
                // with another state, else 0 (-1 at the trailing edge), carried along the row;
                // the elements before the first interior one are exterior
                int lastna = t0 - 1 >= 1 ? t0 - 1 : 0;
                for (int ix = t0; ix < t0 + nt; ix++) {
                    if (rel[ix] < ADHES) {
                        if (ix >= 1) lastna = ix;
                        continue;
                    }
#ifdef GS_PHASE_TIMING
                    const unsigned long long c0 = wall_clock64();
#endif
                    double qsx = 0.0, qsy = 0.0;
                    for (int jx = j0 + lane; jx <= j1; jx += 64) {
                        int o = ix - jx + mx;
                        double qx = rdpx[jx], qy = rdpy[jx];
                        qsx += stc[o] * qx + stc[M2 + o] * qy;
                        qsy += stc[2 * M2 + o] * qx + stc[3 * M2 + o] * qy;
                    }
                    for (int off = 32; off > 0; off >>= 1) {
                        qsx += __shfl_xor(qsx, off);
                        qsy += __shfl_xor(qsy, off);
                    }
#ifdef GS_PHASE_TIMING
                    const unsigned long long c1 = wall_clock64();
#endif
                    const int jx = ix >= 1 ? lastna : -1;
#ifdef GS_PHASE_TIMING
                    const unsigned long long c2 = wall_clock64();
#endif
                    // the element problem, solved by every lane on the same (broadcast) data so that
                    // its change D is known wavefront-wide; lane 0 stores
                    const int8_t eup = ix >= 1 ? rel[ix - 1] : EXTER;
                    double Dx, Dy;
                    {
                        const int b0 = mx, bj = mx + (jx - ix);
                        double a11 = I.ga_inv * (stc[b0] - stc[bj]);
                        double a12 = I.ga_inv * (stc[M2 + b0] - stc[M2 + bj]);
                        double a21 = I.ga_inv * (stc[2 * M2 + b0] - stc[2 * M2 + bj]);
                        double a22 = I.ga_inv * (stc[3 * M2 + b0] - stc[3 * M2 + bj]);
                        double sv[2] = {rwx[ix] + I.ga_inv * (ucx[ix] + qsx), rwy[ix] + I.ga_inv * (ucy[ix] + qsy)};
                        const double opx = rpx[ix], opy = rpy[ix];
                        double pr[3] = {opx, opy, rpn[ix]};
                        int8_t e8 = rel[ix];
                        elmtrc(e8, a11, a12, a21, a22, I.eps, pr, I.mu, sv, I.omegah, I.omegas, true);
                        Dx = pr[0] - opx;
                        Dy = pr[1] - opy;
                        if (lane == 0) {
                            const size_t ii = r0 + ix;
                            difacc += Dx * Dx + Dy * Dy;
                            rdpx[ix] += Dx;
                            rdpy[ix] += Dy;
                            rpx[ix] = pr[0];
                            rpy[ix] = pr[1];
                            rel[ix] = e8;
                            I.sx[ii] = sv[0];
                            I.sy[ii] = sv[1];
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    // Rebuild the upstream part of the row (:3086-3127).  The adhesion elements
                    // upstream satisfy p(k) = p(k+1) + dp(k), so they all move by the change D of
                    // p(ix) until one exceeds its traction bound (clipped there, D becomes that
                    // element's change) or a non-adhesion element ends the chain.  The wavefront
                    // handles 64 upstream elements per step.
#ifdef GS_PHASE_TIMING
                    const unsigned long long c3 = wall_clock64();
#endif
                    int hi = ix - 1;
                    if (ix >= 1 && eup != ADHES) {
                        // the upstream neighbour ends the chain at once
                        if (lane == 0) {
                            const int kx = ix - 1;
                            if (eup == SLIP || eup == PLAST) {
                                rdpx[kx] = rpx[kx] - rpx[ix];
                                rdpy[kx] = rpy[kx] - rpy[ix];
                            } else {
                                rdpx[kx] = -rpx[ix];
                                rdpy[kx] = -rpy[ix];
                            }
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        hi = -1;
                    }
                    while (hi >= 0) {
                        const int kx = hi - lane;
                        const bool valid = kx >= 0;
                        const int8_t ek = valid ? rel[kx] : EXTER;
                        const unsigned long long smask = __ballot(!valid || ek != ADHES);
                        const int sl = smask ? __ffsll((long long)smask) - 1 : 64;
                        double ox = 0.0, oy = 0.0, qx = 0.0, qy = 0.0;
                        bool clip = false;
                        if (lane < sl) {
                            ox = rpx[kx];
                            oy = rpy[kx];
                            qx = ox + Dx;
                            qy = oy + Dy;
                            clip = sqrt(qx * qx + qy * qy) > rg[kx];
                        }
                        const unsigned long long cmask = __ballot(clip);
                        const int cl = cmask ? __ffsll((long long)cmask) - 1 : 64;
                        const int nmove = sl < cl ? sl : cl;
                        if (lane < nmove) {
                            rpx[kx] = qx;
                            rpy[kx] = qy;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        if (cl < sl) {
                            // clip at lane cl and continue below it with that element's change
                            double nDx = 0.0, nDy = 0.0;
                            if (lane == cl) {
                                double ptabs = sqrt(qx * qx + qy * qy), ptbnd = rg[kx];
                                qx = qx * ptbnd / ptabs;
                                qy = qy * ptbnd / ptabs;
                                rdpx[kx] = qx - rpx[kx + 1];
                                rdpy[kx] = qy - rpy[kx + 1];
                                rpx[kx] = qx;
                                rpy[kx] = qy;
                                nDx = qx - ox;
                                nDy = qy - oy;
                            }
                            Dx = __shfl(nDx, cl);
                            Dy = __shfl(nDy, cl);
                            hi -= cl + 1;
                        } else if (sl < 64) {
                            // the first non-adhesion element ends the chain
                            if (lane == sl && valid) {
                                if (ek == SLIP || ek == PLAST) {
                                    rdpx[kx] = rpx[kx] - rpx[kx + 1];
                                    rdpy[kx] = rpy[kx] - rpy[kx + 1];
                                } else if (rel[kx + 1] >= ADHES) {
                                    rdpx[kx] = -rpx[kx + 1];
                                    rdpy[kx] = -rpy[kx + 1];
                                }
                            }
                            hi = -1;
                        } else {
                            hi -= 64;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    }
#ifdef GS_PHASE_TIMING
                    if (lane == 0) {
                        const unsigned long long c4 = wall_clock64();
                        ph[0] += c1 - c0;
                        ph[1] += c2 - c1;
                        ph[2] += c3 - c2;
                        ph[3] += c4 - c3;
                    }
#endif
                    if (ix >= 1 && rel[ix] != ADHES) lastna = ix;
                }
            }
            __syncthreads();
            if (tid == 0) {
                const unsigned long long tr2 = wall_clock64();
                t_rows += tr1 - tr0;
                t_inrow += tr2 - tr1;
            }
            for (int ix = tid; ix < mx; ix += T) {
                I.px[r0 + ix] = rpx[ix];
                I.py[r0 + ix] = rpy[ix];
                dpx[r0 + ix] = rdpx[ix];
                dpy[r0 + ix] = rdpy[ix];
                I.el[r0 + ix] = rel[ix];
            }
            __syncthreads();
        }
        double sq[1] = {0.0};
        if (I.p0x) {
            // sensitivity solve: rms of the change from the initial estimate (:3138-3144)
            for (int i = tid; i < n; i += T) {
                double ax = I.px[i] + -1.0 * I.p0x[i], ay = I.py[i] + -1.0 * I.p0y[i];
                sq[0] += ax * ax + ay * ay;
            }
        } else {
            for (int i = tid; i < n; i += T) sq[0] += I.px[i] * I.px[i] + I.py[i] * I.py[i];
        }
        group_reduce<1>(sq, 0, red_s);
        if (tid == 0) {
            double rms = sqrt(sq[0] / (double)(2 * n > 1 ? 2 * n : 1));
            s_ctl[0] = sqrt(difacc / (2.0 * I.k));
            s_ctl[1] = I.eps * fmax(1e-6, I.facnel * rms);
        }
        __syncthreads();
        dif = s_ctl[0];
        difid = s_ctl[1];
        if (itgs == 1) dif1 = dif;
        __syncthreads();
    }
    if (tid == 0) {
        double len = 0.0;
        for (int jy = 0; jy < my; jy++) len += rj1[jy] >= rj0[jy] ? rj1[jy] - rj0[jy] + 1 : 0;
        ro.partials[(size_t)blockIdx.x * 10] = (double)itgs;
        ro.partials[(size_t)blockIdx.x * 10 + 1] = dif;
        ro.partials[(size_t)blockIdx.x * 10 + 2] = 8.0 * itgs * (double)I.k * len;
        ro.partials[(size_t)blockIdx.x * 10 + 3] = (double)t_rows;
        ro.partials[(size_t)blockIdx.x * 10 + 4] = (double)t_inrow;
        ro.partials[(size_t)blockIdx.x * 10 + 5] = dif1;
        for (int q = 0; q < 4; q++) ro.partials[(size_t)blockIdx.x * 10 + 6 + q] = (double)ph[q];
    }
}

// LDS arrays of the SteadyGS row being walked
struct GsRow {
    double* px; double* py;       // p
    double* dpx; double* dpy;     // differences along x
    double* g;                    // mu pn
    double* pn;
    double* wx; double* wy;       // ws
    double* ucx; double* ucy;     // influence of other rows, unscaled (times ga_inv)
    const double* zx; const double* zy;   // influence of other rows, scaled (null: none)
    const double* stc;            // own-row coefficients, 4 blocks of 2 mx (offsets q - mx)
    int8_t* el;
};


// Wavefront sum on the latency path: butterflies within each row of 16 lanes by DPP moves
// (quad permutes [1,0,3,2] and [2,3,0,1], then the half-row and row mirrors: every lane of a
// row ends with the row's sum, the same bits by commutativity), then the four row sums read
// from lanes 0, 16, 32, 48 and added in that order.  Needs all 64 lanes active.
template <int CTRL>
__device__ inline double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double lane_f64(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ inline double wave_sum(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return ((lane_f64(v, 0) + lane_f64(v, 16)) + lane_f64(v, 32)) + lane_f64(v, 48);
}

// Wavefront 0's walk along elements [ixw, iend) of the row starting at r0 (src/m_solvpt.f90:
// 2950-3127): each interior element's 2x2 problem in order from the trailing edge, then the
// rebuild of the upstream part of the row.  The own-row influence is summed over the
// summation range [j0, j1].  lastna carries the first upstream element not in adhesion along
// the row; difacc (lane 0) sums the squared changes; gsx, gsy receive s of each element.
// Called by all 64 lanes of wavefront 0.
__device__ void gs_walk_span(const GsRow& R, int mx, int j0, int j1, size_t r0, int& ixw, int iend,
                             int& lastna, double& difacc, double ga_inv, double eps, double mu,
                             double omegah, double omegas, double* gsx, double* gsy,
                             bool s_lds = false)
{
    const int lane = threadIdx.x, M2 = 2 * mx;
    for (; ixw < iend; ixw++) {
        const int ix = ixw;
        if (R.el[ix] < ADHES) {
            if (ix >= 1) lastna = ix;
            continue;
        }
        double qsx = 0.0, qsy = 0.0;
        for (int jx = j0 + lane; jx <= j1; jx += 64) {
            int o = ix - jx + mx;
            double qx = R.dpx[jx], qy = R.dpy[jx];
            qsx += R.stc[o] * qx + R.stc[M2 + o] * qy;
            qsy += R.stc[2 * M2 + o] * qx + R.stc[3 * M2 + o] * qy;
        }
        qsx = wave_sum(qsx);
        qsy = wave_sum(qsy);
        const int jx = ix >= 1 ? lastna : -1;
        // the element problem, solved by every lane on the same (broadcast) data so
        // that its change D is known wavefront-wide; lane 0 stores
        const int8_t eup = ix >= 1 ? R.el[ix - 1] : EXTER;
        double Dx, Dy;
        {
            const int b0 = mx, bj = mx + (jx - ix);
            double a11 = ga_inv * (R.stc[b0] - R.stc[bj]);
            double a12 = ga_inv * (R.stc[M2 + b0] - R.stc[M2 + bj]);
            double a21 = ga_inv * (R.stc[2 * M2 + b0] - R.stc[2 * M2 + bj]);
            double a22 = ga_inv * (R.stc[3 * M2 + b0] - R.stc[3 * M2 + bj]);
            double sv[2];
            if (R.zx) {
                sv[0] = (R.wx[ix] + R.zx[ix]) + ga_inv * (R.ucx[ix] + qsx);
                sv[1] = (R.wy[ix] + R.zy[ix]) + ga_inv * (R.ucy[ix] + qsy);
            } else {
                sv[0] = R.wx[ix] + ga_inv * (R.ucx[ix] + qsx);
                sv[1] = R.wy[ix] + ga_inv * (R.ucy[ix] + qsy);
            }
            const double opx = R.px[ix], opy = R.py[ix];
            double pr[3] = {opx, opy, R.pn[ix]};
            int8_t e8 = R.el[ix];
            elmtrc(e8, a11, a12, a21, a22, eps, pr, mu, sv, omegah, omegas, true);
            Dx = pr[0] - opx;
            Dy = pr[1] - opy;
            if (lane == 0) {
                const size_t ii = r0 + ix;
                difacc += Dx * Dx + Dy * Dy;
                R.dpx[ix] += Dx;
                R.dpy[ix] += Dy;
                R.px[ix] = pr[0];
                R.py[ix] = pr[1];
                R.el[ix] = e8;
                if (s_lds) {
                    // kept in LDS in place of this element's other-row influence (read once,
                    // above) and stored with the row: no global store on the element chain
                    R.ucx[ix] = sv[0];
                    R.ucy[ix] = sv[1];
                } else {
                    gsx[ii] = sv[0];
                    gsy[ii] = sv[1];
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Rebuild the upstream part of the row (:3086-3127).  The adhesion elements
        // upstream satisfy p(k) = p(k+1) + dp(k), so they all move by the change D
        // of p(ix) until one exceeds its traction bound (clipped there, D becomes
        // that element's change) or a non-adhesion element ends the chain.  The
        // wavefront handles 64 upstream elements per step.
        int hi = ix - 1;
        if (ix >= 1 && eup != ADHES) {
            // the upstream neighbour ends the chain at once
            if (lane == 0) {
                const int kx = ix - 1;
                if (eup == SLIP || eup == PLAST) {
                    R.dpx[kx] = R.px[kx] - R.px[ix];
                    R.dpy[kx] = R.py[kx] - R.py[ix];
                } else {
                    R.dpx[kx] = -R.px[ix];
                    R.dpy[kx] = -R.py[ix];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            hi = -1;
        }
        while (hi >= 0) {
            const int kx = hi - lane;
            const bool valid = kx >= 0;
            const int8_t ek = valid ? R.el[kx] : EXTER;
            const unsigned long long smask = __ballot(!valid || ek != ADHES);
            const int sl = smask ? __ffsll((long long)smask) - 1 : 64;
            double ox = 0.0, oy = 0.0, qx = 0.0, qy = 0.0;
            bool clip = false;
            if (lane < sl) {
                ox = R.px[kx];
                oy = R.py[kx];
                qx = ox + Dx;
                qy = oy + Dy;
                clip = sqrt(qx * qx + qy * qy) > R.g[kx];
            }
            const unsigned long long cmask = __ballot(clip);
            const int cl = cmask ? __ffsll((long long)cmask) - 1 : 64;
            const int nmove = sl < cl ? sl : cl;
            if (lane < nmove) {
                R.px[kx] = qx;
                R.py[kx] = qy;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (cl < sl) {
                // clip at lane cl and continue below it with that element's change
                double nDx = 0.0, nDy = 0.0;
                if (lane == cl) {
                    double ptabs = sqrt(qx * qx + qy * qy), ptbnd = R.g[kx];
                    qx = qx * ptbnd / ptabs;
                    qy = qy * ptbnd / ptabs;
                    R.dpx[kx] = qx - R.px[kx + 1];
                    R.dpy[kx] = qy - R.py[kx + 1];
                    R.px[kx] = qx;
                    R.py[kx] = qy;
                    nDx = qx - ox;
                    nDy = qy - oy;
                }
                Dx = __shfl(nDx, cl);
                Dy = __shfl(nDy, cl);
                hi -= cl + 1;
            } else if (sl < 64) {
                // the first non-adhesion element ends the chain
                if (lane == sl && valid) {
                    if (ek == SLIP || ek == PLAST) {
                        R.dpx[kx] = R.px[kx] - R.px[kx + 1];
                        R.dpy[kx] = R.py[kx] - R.py[kx + 1];
                    } else if (R.el[kx + 1] >= ADHES) {
                        R.dpx[kx] = -R.px[kx + 1];
                        R.dpy[kx] = -R.py[kx + 1];
                    }
                }
                hi = -1;
            } else {
                hi -= 64;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (ix >= 1 && R.el[ix] != ADHES) lastna = ix;
    }
}

// SteadyGS with the row sums of the next row overlapped with the walk along the current row:
// while wavefront 0 walks row iy element by element, the other wavefronts sum the influence
// of every source row but iy and iy' on the next target row iy' (those rows do not change
// while row iy is walked); row iy's own contribution to row iy' is added once its walk has
// ended.  Same sweeps and element updates as k_stdygs2; only the summation order of the
// row sums differs.  Returns the same values as k_stdygs2, with the ticks of the row-sum
// phases and of wavefront 0's walks in place of the two phase times.
__host__ __device__ inline size_t gs3_base_lds(int mx, int my)
{
    return (size_t)(18 * mx) * sizeof(double) + 2 * GS_THREADS * sizeof(double) +
           (size_t)4 * my * sizeof(int) + mx + 64;
}
__host__ __device__ inline size_t gs3_stage_off(int mx, int my)
{
    return (gs3_base_lds(mx, my) + 63) / 64 * 64;
}
// a helper thread owns GS3_TB consecutive targets and takes the sources in chunks of GS3_QB
// (4 x 8: as many products per coefficient load as 8 x 4, with half the accumulators, which
// stay live while wavefront 0 walks)
#ifndef GS3_TB_N
#define GS3_TB_N 4
#endif
#ifndef GS3_QB_N
#define GS3_QB_N 8
#endif
constexpr int GS3_TB = GS3_TB_N, GS3_QB = GS3_QB_N;
__host__ __device__ inline int gs3_window(int mx) { return (2 * mx + GS3_TB + GS3_QB + 3) / 4 * 4; }
// one staging buffer: room for four coefficient windows (the row sums use three: xx, xy, yy)
// and the two difference rows
__host__ __device__ inline int gs3_stage_stride(int mx)
{
    return (4 * gs3_window(mx) + 2 * (mx + GS3_QB) + 7) / 8 * 8;
}
__host__ __device__ inline size_t gs3_lds(int mx, int my)
{
    return gs3_stage_off(mx, my) + 2 * (size_t)gs3_stage_stride(mx) * sizeof(double);
}

__global__ void __launch_bounds__(GS_THREADS) k_stdygs3(const GsItem* items, const int* gstart,
                                                        int nitems, RedOut ro)
{
    extern __shared__ double gs_lds[];
    __shared__ double red_s[16 * MAXK];
    __shared__ double s_ctl[4];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsItem& I = items[it];
    const int mx = I.mx, my = I.my, n = I.n, tid = threadIdx.x, T = blockDim.x;
    const int M2 = 2 * mx;
    // helpers: the wavefronts after the first
    const int TH = T - 64, hid = tid - 64;
    const bool helper = tid >= 64;
    double* rpx = gs_lds;
    double* rpy = rpx + mx;
    double* rdpx = rpy + mx;
    double* rdpy = rdpx + mx;
    double* rg = rdpy + mx;
    double* rpn = rg + mx;
    double* rwx = rpn + mx;
    double* rwy = rwx + mx;
    double* ucx = rwy + mx;                // sums over the other rows, per column
    double* ucy = ucx + mx;
    double* stc = ucy + mx;                // own-row coefficients, offsets (q - mx, 0)
    double* red = stc + 4 * M2;            // 2 partial sums per thread
    int* rj0 = (int*)(red + 2 * GS_THREADS);   // summation range per row
    int* rj1 = rj0 + my;
    int* ri0 = rj1 + my;                   // interior range per row
    int* ri1 = ri0 + my;
    int8_t* rel = (int8_t*)(ri1 + my);
    double* stg = gs_lds + gs3_stage_off(mx, my) / 8;
    const int SW = gs3_stage_stride(mx);
    const GsRow R{rpx, rpy, rdpx, rdpy, rg, rpn, rwx, rwy, ucx, ucy, nullptr, nullptr, stc, rel};
    double* dpx = I.dpx;
    double* dpy = I.dpy;
    // the xx, xy and yy coefficient blocks: the coefficient kernel writes the xy and yx blocks
    // from one value, so the yx block is the xy block
    const double* cfs[3] = {I.cf[0][0], I.cf[0][1], I.cf[1][1]};
    const double* cfb[4] = {I.cf[0][0], I.cf[0][1], I.cf[1][0], I.cf[1][1]};

    for (int i = tid; i < n; i += T) {
        int ix = i % mx;
        if (ix != mx - 1) {
            dpx[i] = I.px[i] - I.px[i + 1];
            dpy[i] = I.py[i] - I.py[i + 1];
        } else {
            dpx[i] = I.px[i];
            dpy[i] = I.py[i];
        }
    }
    for (int jy = tid; jy < my; jy += T) {
        const int8_t* row = I.el + (size_t)jy * mx;
        int first = 0;
        while (first < mx - 1 && row[first] <= EXTER) first++;
        int last = -1, lo = mx;
        for (int ix = 0; ix < mx; ix++)
            if (row[ix] >= ADHES) {
                last = ix;
                if (ix < lo) lo = ix;
            }
        rj0[jy] = first - 1 < 0 ? 0 : first - 1;
        rj1[jy] = last + 1 > mx - 1 ? mx - 1 : last + 1;
        ri0[jy] = lo;
        ri1[jy] = last;
    }
    for (int e = tid; e < 4 * M2; e += T) {
        int b = e / M2, q = e - b * M2;
        stc[e] = cfb[b][(size_t)my * M2 + q];
    }
    __syncthreads();
    int nsrc = 0;   // source rows
    for (int jy = 0; jy < my; jy++) nsrc += rj1[jy] >= rj0[jy] ? 1 : 0;

    // ---- the row sums of target row tr over the source rows other than tr and skip, into the
    // helpers' accumulators; wavefront 0 meanwhile walks its row (walk(b, nb): the share of
    // the row due after b of nb source rows)
    double ax[GS3_TB], ay[GS3_TB];
    int sgrp = 0;
    auto row_sums = [&](int tr, int skip, int only, auto&& walk) {
        const int t0 = ri0[tr], nt = ri1[tr] - t0 + 1;
        const int ntg = (nt + GS3_TB - 1) / GS3_TB, ntp = ntg * GS3_TB;
        const int P = TH / ntg > 0 ? TH / ntg : 1;
        const int tg = hid % ntg, part = hid / ntg;
        const bool active = helper && part < P;
        const int nb = only >= 0 ? 1 : nsrc - (rj1[tr] >= rj0[tr]) - (skip >= 0 && rj1[skip] >= rj0[skip]);
        int jy = 0, b = 0;
        while (true) {
            int sy = -1;
            if (only >= 0) {
                if (b == 0) sy = only;
            } else {
                for (; jy < my; jy++)
                    if (jy != tr && jy != skip && rj1[jy] >= rj0[jy]) {
                        sy = jy++;
                        break;
                    }
            }
            if (sy < 0) break;
            double* sc = stg + (size_t)sgrp * SW;
            sgrp ^= 1;
            const int j0 = rj0[sy], nq = rj1[sy] - j0 + 1;
            const int Q = (nq + GS3_QB - 1) / GS3_QB * GS3_QB;
            const int Wn = ntp + Q - 1, Wp = (Wn + 1 + 3) / 4 * 4;
            if (helper) {
                // window S[k], k = t - q + Q - 1: coefficient of target t0 + t and source j0 + q
                const int col0 = t0 - j0 - (Q - 1) + mx;
#pragma unroll
                for (int bb = 0; bb < 3; bb++) {
                    const double* crow = cfs[bb] + (size_t)(tr - sy + my) * M2;
                    for (int k = hid; k < Wp; k += TH) {
                        const int col = col0 + k;
                        sc[bb * Wp + k] = (k < Wn && col >= 0 && col < M2) ? crow[col] : 0.0;
                    }
                }
                double* sdx = sc + 3 * Wp;
                double* sdy = sdx + Q;
                for (int e = hid; e < Q; e += TH) {
                    const bool in = e < nq;
                    sdx[e] = in ? dpx[(size_t)sy * mx + j0 + e] : 0.0;
                    sdy[e] = in ? dpy[(size_t)sy * mx + j0 + e] : 0.0;
                }
            } else {
                walk(b, nb);
            }
            __syncthreads();
            b++;
            if (active) {
                const int nch = Q / GS3_QB;
                const double* sdx = sc + 3 * Wp;
                const double* sdy = sdx + Q;
                for (int c = part; c < nch; c += P) {
                    const int kmin = tg * GS3_TB - c * GS3_QB + (Q - GS3_QB);
                    double qx[GS3_QB], qy[GS3_QB];
#pragma unroll
                    for (int j = 0; j < GS3_QB; j++) {
                        qx[j] = sdx[c * GS3_QB + j];
                        qy[j] = sdy[c * GS3_QB + j];
                    }
                    // windows of the xx, xy (= yx) and yy blocks
                    const double2* w0 = (const double2*)(sc + kmin);
                    const double2* w1 = (const double2*)(sc + Wp + kmin);
                    const double2* w3 = (const double2*)(sc + 2 * Wp + kmin);
                    double v0[GS3_TB + GS3_QB], v1[GS3_TB + GS3_QB], v3[GS3_TB + GS3_QB];
#pragma unroll
                    for (int m = 0; m < (GS3_TB + GS3_QB) / 2; m++) {
                        const double2 a = w0[m], b2 = w1[m], d = w3[m];
                        v0[2 * m] = a.x;
                        v0[2 * m + 1] = a.y;
                        v1[2 * m] = b2.x;
                        v1[2 * m + 1] = b2.y;
                        v3[2 * m] = d.x;
                        v3[2 * m + 1] = d.y;
                    }
#pragma unroll
                    for (int i = 0; i < GS3_TB; i++) {
                        double accx = ax[i], accy = ay[i];
#pragma unroll
                        for (int j = 0; j < GS3_QB; j++) {
                            const int m = i + GS3_QB - 1 - j;
                            accx += v0[m] * qx[j] + v1[m] * qy[j];
                            accy += v1[m] * qx[j] + v3[m] * qy[j];
                        }
                        ax[i] = accx;
                        ay[i] = accy;
                    }
                }
            }
        }
        if (!helper) walk(nb, nb);
    };
    // sums over the helpers' parts (target row tr)
    auto reduce_sums = [&](int tr) {
        const int t0 = ri0[tr], nt = ri1[tr] - t0 + 1;
        const int ntg = (nt + GS3_TB - 1) / GS3_TB;
        const int P = TH / ntg > 0 ? TH / ntg : 1;
        const int tg = hid % ntg, part = hid / ntg;
        const bool active = helper && part < P;
        __syncthreads();
        for (int i = 0; i < GS3_TB; i++) {
            if (active) {
                red[2 * (part * ntg + tg)] = ax[i];
                red[2 * (part * ntg + tg) + 1] = ay[i];
            }
            __syncthreads();
            for (int g = tid; g < ntg; g += T) {
                const int c = g * GS3_TB + i;
                if (c < nt) {
                    double sx = 0.0, sy = 0.0;
                    for (int p = 0; p < P; p++) {
                        sx += red[2 * (p * ntg + g)];
                        sy += red[2 * (p * ntg + g) + 1];
                    }
                    ucx[t0 + c] = sx;
                    ucy[t0 + c] = sy;
                }
            }
            __syncthreads();
        }
    };
    auto clear_acc = [&]() {
#pragma unroll
        for (int i = 0; i < GS3_TB; i++) ax[i] = ay[i] = 0.0;
    };
    auto load_row = [&](int iy) {
        const size_t r0 = (size_t)iy * mx;
        for (int ix = tid; ix < mx; ix += T) {
            rpx[ix] = I.px[r0 + ix];
            rpy[ix] = I.py[r0 + ix];
            rdpx[ix] = dpx[r0 + ix];
            rdpy[ix] = dpy[r0 + ix];
            double pn = I.pn[r0 + ix];
            rpn[ix] = pn;
            rg[ix] = I.mu * pn;
            rwx[ix] = I.wsx[r0 + ix];
            rwy[ix] = I.wsy[r0 + ix];
            rel[ix] = I.el[r0 + ix];
        }
    };
    auto store_row = [&](int iy) {
        const size_t r0 = (size_t)iy * mx;
        for (int ix = tid; ix < mx; ix += T) {
            I.px[r0 + ix] = rpx[ix];
            I.py[r0 + ix] = rpy[ix];
            dpx[r0 + ix] = rdpx[ix];
            dpy[r0 + ix] = rdpy[ix];
            I.el[r0 + ix] = rel[ix];
        }
    };
    auto next_row = [&](int iy) {
        int r = iy + 1;
        while (r < my && ri1[r] < ri0[r]) r++;
        return r;
    };

    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0;
    unsigned long long t_rows = 0, t_walk = 0;   // thread 0: wall-clock ticks
    auto nowalk = [&](int, int) {};
    while (dif >= difid && itgs < I.maxgs) {
        itgs++;
        double difacc = 0.0;   // thread 0
        int iy = next_row(-1);
        if (iy < my) {
            const unsigned long long tr0 = wall_clock64();
            clear_acc();
            row_sums(iy, -1, -1, nowalk);
            reduce_sums(iy);
            load_row(iy);
            __syncthreads();
            if (tid == 0) t_rows += wall_clock64() - tr0;
        }
        while (iy < my) {
            const int nx = next_row(iy);
            // the walk along row iy by wavefront 0 (one element at a time)
            const int t0 = ri0[iy], nt = ri1[iy] - t0 + 1;
            const int j0 = rj0[iy], j1 = rj1[iy];
            const size_t r0 = (size_t)iy * mx;
            const int lane = tid;
            // first upstream element not in adhesion (:2972-2977): the largest jx in [1, ix-1]
            // with another state, else 0 (-1 at the trailing edge), carried along the row;
            // the elements before the first interior one are exterior
            int lastna = t0 - 1 >= 1 ? t0 - 1 : 0;
            int ixw = t0;
            auto walk = [&](int b, int nb) {
                const int iend = b >= nb ? t0 + nt : t0 + (int)(((long)nt * (b + 1) + nb - 1) / nb);
                if (ixw >= iend) return;
                const unsigned long long w0 = wall_clock64();
                gs_walk_span(R, mx, j0, j1, r0, ixw, iend, lastna, difacc, I.ga_inv, I.eps, I.mu,
                             I.omegah, I.omegas, I.sx, I.sy);
                if (lane == 0) t_walk += wall_clock64() - w0;
            };
            const unsigned long long tr0 = wall_clock64();
            if (nx < my) {
                // the next row's sums over the rows that do not change meanwhile
                clear_acc();
                row_sums(nx, iy, -1, walk);
            } else if (!helper) {
                walk(1, 1);
            }
            __syncthreads();
            store_row(iy);
            __syncthreads();
            if (nx < my) {
                // ... and row iy's contribution, now final
                if (rj1[iy] >= rj0[iy]) row_sums(nx, -1, iy, nowalk);
                reduce_sums(nx);
                load_row(nx);
                __syncthreads();
            }
            if (tid == 0) t_rows += wall_clock64() - tr0;
            iy = nx;
        }
        double sq[1] = {0.0};
        if (I.p0x) {
            // sensitivity solve: rms of the change from the initial estimate (:3138-3144)
            for (int i = tid; i < n; i += T) {
                double bx = I.px[i] + -1.0 * I.p0x[i], by = I.py[i] + -1.0 * I.p0y[i];
                sq[0] += bx * bx + by * by;
            }
        } else {
            for (int i = tid; i < n; i += T) sq[0] += I.px[i] * I.px[i] + I.py[i] * I.py[i];
        }
        group_reduce<1>(sq, 0, red_s);
        if (tid == 0) {
            double rms = sqrt(sq[0] / (double)(2 * n > 1 ? 2 * n : 1));
            s_ctl[0] = sqrt(difacc / (2.0 * I.k));
            s_ctl[1] = I.eps * fmax(1e-6, I.facnel * rms);
        }
        __syncthreads();
        dif = s_ctl[0];
        difid = s_ctl[1];
        if (itgs == 1) dif1 = dif;
        __syncthreads();
    }
    if (tid == 0) {
        double len = 0.0;
        for (int jy = 0; jy < my; jy++) len += rj1[jy] >= rj0[jy] ? rj1[jy] - rj0[jy] + 1 : 0;
        ro.partials[(size_t)blockIdx.x * 10] = (double)itgs;
        ro.partials[(size_t)blockIdx.x * 10 + 1] = dif;
        ro.partials[(size_t)blockIdx.x * 10 + 2] = 8.0 * itgs * (double)I.k * len;
        ro.partials[(size_t)blockIdx.x * 10 + 3] = (double)t_rows;
        ro.partials[(size_t)blockIdx.x * 10 + 4] = (double)t_walk;
        ro.partials[(size_t)blockIdx.x * 10 + 5] = dif1;
        for (int q = 6; q < 10; q++) ro.partials[(size_t)blockIdx.x * 10 + q] = 0.0;
    }
}

// ------------------------------------------------------------------------------------------
// SteadyGS by blocks of rows (GsbItem): set-up, the walk of one block, the end-of-sweep test
// ------------------------------------------------------------------------------------------

constexpr int GSB_THREADS = 256;

// set-up: dp = differences along x (the leading-edge column keeps p, src/m_solvpt.f90:
// 2893-2908), the summation and interior range of each row, d = 0 and the sweep state.
// Returns (first interior row, last interior row, summed range lengths).
__global__ void __launch_bounds__(GSB_THREADS) k_gsb_setup(const GsbItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    __shared__ double red[16 * MAXK];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsbItem& I = items[it];
    const int mx = I.mx, my = I.my, n = I.n, tid = threadIdx.x, T = blockDim.x;
    for (int i = tid; i < n; i += T) {
        const int ix = i % mx;
        if (ix != mx - 1) {
            I.dpx[i] = I.px[i] - I.px[i + 1];
            I.dpy[i] = I.py[i] - I.py[i + 1];
        } else {
            I.dpx[i] = I.px[i];
            I.dpy[i] = I.py[i];
        }
        I.ddx[i] = 0.0;
        I.ddy[i] = 0.0;
    }
    // summation range [first - 1, last + 1] (src/m_aijpj.f90:173-185) and interior range
    double acc[3] = {INFINITY, -INFINITY, 0.0};   // -first row, last row, lengths
    for (int jy = tid; jy < my; jy += T) {
        const int8_t* row = I.el + (size_t)jy * mx;
        int first = 0;
        while (first < mx - 1 && row[first] <= EXTER) first++;
        int last = -1, lo = mx;
        for (int ix = 0; ix < mx; ix++)
            if (row[ix] >= ADHES) {
                last = ix;
                if (ix < lo) lo = ix;
            }
        const int j0 = first - 1 < 0 ? 0 : first - 1, j1 = last + 1 > mx - 1 ? mx - 1 : last + 1;
        I.rng[jy] = j0;
        I.rng[my + jy] = j1;
        I.rng[2 * my + jy] = lo;
        I.rng[3 * my + jy] = last;
        if (last >= lo) {
            acc[0] = fmin(acc[0], (double)jy);
            acc[1] = fmax(acc[1], (double)jy);
        }
        if (j1 >= j0) acc[2] += (double)(j1 - j0 + 1);
    }
    // (the three values reduce with min, max and sum)
    double a0[1] = {acc[0]}, a1[1] = {acc[1]}, a2[1] = {acc[2]};
    group_reduce<1>(a0, 2, red);
    group_reduce<1>(a1, 1, red);
    group_reduce<1>(a2, 0, red);
    if (tid == 0) {
        for (int q = 0; q < 4; q++) I.st[q] = 0.0;
        double* r = ro.partials + (size_t)blockIdx.x * 10;
        r[0] = a0[0];
        r[1] = a1[0];
        r[2] = a2[0];
        for (int q = 3; q < 10; q++) r[q] = 0.0;
    }
}

// LDS of k_gsb_walk: 14 row arrays, the own-row coefficients (4 x 2 mx), the staging of a
// row product (3 coefficient windows and 2 source rows), the states of the row
__host__ __device__ inline int gsb_window(int mx) { return (2 * mx + GS3_TB + GS3_QB + 3) / 4 * 4; }
__host__ __device__ inline size_t gsb_lds(int mx)
{
    return ((size_t)14 * mx + 8 * mx + 3 * gsb_window(mx) + 2 * (mx + GS3_QB) + 2 * GSB_THREADS) *
               sizeof(double) + mx + 64;
}

// The walk of one block of rows.  For row t of the block, the influence of every other row is
//   z_t - A(0) dp_t + sum over the block's rows j < t of A(t - j) d_j
// (z = A dp at the block's start includes row t itself and the block's rows at their
// start values; d_j = the change of row j's differences in this block), then wavefront 0
// walks the row with the own-row influence of the current differences, as k_stdygs3 does.
// The row products run on all threads: a thread owns GS3_TB consecutive targets and takes
// the sources in chunks of GS3_QB from one coefficient window per block (xx, xy = yx, yy).
__global__ void __launch_bounds__(GSB_THREADS) k_gsb_walk(const GsbItem* items, const int* gstart,
                                                          int nitems, RedOut ro)
{
    extern __shared__ double gsb_lds_[];
    (void)ro;
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsbItem& I = items[it];
    const int mx = I.mx, my = I.my, tid = threadIdx.x, T = blockDim.x, M2 = 2 * mx;
    double* rpx = gsb_lds_;
    double* rpy = rpx + mx;
    double* rdpx = rpy + mx;
    double* rdpy = rdpx + mx;
    double* r0x = rdpy + mx;               // differences at the row's start
    double* r0y = r0x + mx;
    double* rg = r0y + mx;
    double* rpn = rg + mx;
    double* rwx = rpn + mx;
    double* rwy = rwx + mx;
    double* ucx = rwy + mx;
    double* ucy = ucx + mx;
    double* rzx = ucy + mx;
    double* rzy = rzx + mx;
    double* stc = rzy + mx;                // own-row coefficients, offsets (q - mx, 0)
    double* sc = stc + 4 * M2;             // product staging: 3 windows, then 2 source rows
    double* red = sc + 3 * gsb_window(mx) + 2 * (mx + GS3_QB);
    int8_t* rel = (int8_t*)(red + 2 * GSB_THREADS);
    const int* rj0 = I.rng;
    const int* rj1 = rj0 + my;
    const int* ri0 = rj1 + my;
    const int* ri1 = ri0 + my;
    const double* cfs[3] = {I.cf[0][0], I.cf[0][1], I.cf[1][1]};   // yx = xy
    const double* cfb[4] = {I.cf[0][0], I.cf[0][1], I.cf[1][0], I.cf[1][1]};
    const GsRow R{rpx, rpy, rdpx, rdpy, rg, rpn, rwx, rwy, ucx, ucy, rzx, rzy, stc, rel};

    // d of the previous block has been used by its product: clear it
    for (size_t e = (size_t)I.z0 * mx + tid; e < (size_t)I.z1 * mx; e += T) {
        I.ddx[e] = 0.0;
        I.ddy[e] = 0.0;
    }
    for (int e = tid; e < 4 * M2; e += T) {
        const int b = e / M2, q = e - b * M2;
        stc[e] = cfb[b][(size_t)my * M2 + q];
    }
    double difacc = 0.0, flops = 0.0;   // thread 0
    double ax[GS3_TB], ay[GS3_TB];
    // ax, ay (targets of row tr) += A(tr - sy) src over the source range [j0, j0 + nq), the
    // source row given in global memory (sgx, sgy) or, for the row itself, in LDS negated
    auto row_prod = [&](int tr, int sy, const double* sgx, const double* sgy, bool self) {
        const int t0 = ri0[tr], nt = ri1[tr] - t0 + 1;
        const int j0 = rj0[sy], nq = rj1[sy] - j0 + 1;
        if (nq <= 0) return;
        const int ntg = (nt + GS3_TB - 1) / GS3_TB, ntp = ntg * GS3_TB;
        const int P = T / ntg > 0 ? T / ntg : 1;
        const int tg = tid % ntg, part = tid / ntg;
        const int Q = (nq + GS3_QB - 1) / GS3_QB * GS3_QB;
        const int Wn = ntp + Q - 1, Wp = (Wn + 1 + 3) / 4 * 4;
        // window S[k], k = t - q + Q - 1: coefficient of target t0 + t and source j0 + q
        const int col0 = t0 - j0 - (Q - 1) + mx;
        __syncthreads();
#pragma unroll
        for (int bb = 0; bb < 3; bb++) {
            const double* crow = cfs[bb] + (size_t)(tr - sy + my) * M2;
            for (int k = tid; k < Wp; k += T) {
                const int col = col0 + k;
                sc[bb * Wp + k] = (k < Wn && col >= 0 && col < M2) ? crow[col] : 0.0;
            }
        }
        double* sdx = sc + 3 * Wp;
        double* sdy = sdx + Q;
        for (int e = tid; e < Q; e += T) {
            const bool in = e < nq;
            if (self) {
                sdx[e] = in ? -r0x[j0 + e] : 0.0;
                sdy[e] = in ? -r0y[j0 + e] : 0.0;
            } else {
                sdx[e] = in ? sgx[(size_t)sy * mx + j0 + e] : 0.0;
                sdy[e] = in ? sgy[(size_t)sy * mx + j0 + e] : 0.0;
            }
        }
        __syncthreads();
        if (part < P) {
            const int nch = Q / GS3_QB;
            for (int c = part; c < nch; c += P) {
                const int kmin = tg * GS3_TB - c * GS3_QB + (Q - GS3_QB);
                double qx[GS3_QB], qy[GS3_QB];
#pragma unroll
                for (int j = 0; j < GS3_QB; j++) {
                    qx[j] = sdx[c * GS3_QB + j];
                    qy[j] = sdy[c * GS3_QB + j];
                }
                const double2* w0 = (const double2*)(sc + kmin);
                const double2* w1 = (const double2*)(sc + Wp + kmin);
                const double2* w3 = (const double2*)(sc + 2 * Wp + kmin);
                double v0[GS3_TB + GS3_QB], v1[GS3_TB + GS3_QB], v3[GS3_TB + GS3_QB];
#pragma unroll
                for (int m = 0; m < (GS3_TB + GS3_QB) / 2; m++) {
                    const double2 a = w0[m], b2 = w1[m], d = w3[m];
                    v0[2 * m] = a.x;
                    v0[2 * m + 1] = a.y;
                    v1[2 * m] = b2.x;
                    v1[2 * m + 1] = b2.y;
                    v3[2 * m] = d.x;
                    v3[2 * m + 1] = d.y;
                }
#pragma unroll
                for (int i = 0; i < GS3_TB; i++) {
                    double accx = ax[i], accy = ay[i];
#pragma unroll
                    for (int j = 0; j < GS3_QB; j++) {
                        const int m = i + GS3_QB - 1 - j;
                        accx += v0[m] * qx[j] + v1[m] * qy[j];
                        accy += v1[m] * qx[j] + v3[m] * qy[j];
                    }
                    ax[i] = accx;
                    ay[i] = accy;
                }
            }
        }
        if (tid == 0) flops += 8.0 * nt * nq;
    };
    // ucx, ucy over the interior of row tr = the sum of the parts' accumulators
    auto reduce_parts = [&](int tr) {
        const int t0 = ri0[tr], nt = ri1[tr] - t0 + 1;
        const int ntg = (nt + GS3_TB - 1) / GS3_TB;
        const int P = T / ntg > 0 ? T / ntg : 1;
        const int tg = tid % ntg, part = tid / ntg;
        __syncthreads();
        for (int i = 0; i < GS3_TB; i++) {
            if (part < P) {
                red[2 * (part * ntg + tg)] = ax[i];
                red[2 * (part * ntg + tg) + 1] = ay[i];
            }
            __syncthreads();
            for (int g = tid; g < ntg; g += T) {
                const int c = g * GS3_TB + i;
                if (c < nt) {
                    double sx = 0.0, sy = 0.0;
                    for (int p = 0; p < P; p++) {
                        sx += red[2 * (p * ntg + g)];
                        sy += red[2 * (p * ntg + g) + 1];
                    }
                    ucx[t0 + c] = sx;
                    ucy[t0 + c] = sy;
                }
            }
            __syncthreads();
        }
    };

    for (int t = I.y0; t < I.y1; t++) {
        if (ri1[t] < ri0[t]) continue;   // no interior element in this row
        const size_t r0 = (size_t)t * mx;
        __syncthreads();
        for (int ix = tid; ix < mx; ix += T) {
            rpx[ix] = I.px[r0 + ix];
            rpy[ix] = I.py[r0 + ix];
            const double dx = I.dpx[r0 + ix], dy = I.dpy[r0 + ix];
            rdpx[ix] = dx;
            rdpy[ix] = dy;
            r0x[ix] = dx;
            r0y[ix] = dy;
            const double pn = I.pn[r0 + ix];
            rpn[ix] = pn;
            rg[ix] = I.mu * pn;
            rwx[ix] = I.wsx[r0 + ix];
            rwy[ix] = I.wsy[r0 + ix];
            rzx[ix] = I.zx[r0 + ix];
            rzy[ix] = I.zy[r0 + ix];
            rel[ix] = I.el[r0 + ix];
        }
#pragma unroll
        for (int i = 0; i < GS3_TB; i++) ax[i] = ay[i] = 0.0;
        // minus the row's own influence at its start (z holds it), plus the changes of the
        // block's rows already walked
        row_prod(t, t, nullptr, nullptr, true);
        for (int j = I.y0; j < t; j++)
            if (rj1[j] >= rj0[j]) row_prod(t, j, I.ddx, I.ddy, false);
        reduce_parts(t);
        if (tid < 64) {
            const int t0 = ri0[t], nt = ri1[t] - t0 + 1;
            int lastna = t0 - 1 >= 1 ? t0 - 1 : 0;
            int ixw = t0;
            gs_walk_span(R, mx, rj0[t], rj1[t], r0, ixw, t0 + nt, lastna, difacc, I.ga_inv, I.eps,
                         I.mu, I.omegah, I.omegas, I.sx, I.sy, true);
        }
        __syncthreads();
        const int t1 = ri1[t];
        for (int ix = tid; ix < mx; ix += T) {
            if (ix >= ri0[t] && ix <= t1 && rel[ix] >= ADHES) {
                I.sx[r0 + ix] = ucx[ix];
                I.sy[r0 + ix] = ucy[ix];
            }
            I.px[r0 + ix] = rpx[ix];
            I.py[r0 + ix] = rpy[ix];
            I.dpx[r0 + ix] = rdpx[ix];
            I.dpy[r0 + ix] = rdpy[ix];
            I.ddx[r0 + ix] = rdpx[ix] - r0x[ix];
            I.ddy[r0 + ix] = rdpy[ix] - r0y[ix];
            I.el[r0 + ix] = rel[ix];
        }
    }
    if (tid == 0) {
        I.st[0] += difacc;
        I.st[3] += flops;
    }
}

// end of a sweep (src/m_solvpt.f90:3130-3150): dif = sqrt(sum |dp|^2 / 2k), difid from the rms
// of p (of p - p0 in a sensitivity solve); counts the sweep and clears its sum.  Returns
// (itgs, dif, flops, 0, 0, dif1, still sweeping, difid, 0, 0).
__global__ void __launch_bounds__(GSB_THREADS) k_gsb_check(const GsbItem* items, const int* gstart,
                                                           int nitems, RedOut ro)
{
    __shared__ double red[16 * MAXK];
    int it = find_item(gstart, nitems, blockIdx.x);
    const GsbItem& I = items[it];
    const int n = I.n, tid = threadIdx.x, T = blockDim.x;
    double sq[1] = {0.0};
    if (I.p0x) {
        for (int i = tid; i < n; i += T) {
            double bx = I.px[i] + -1.0 * I.p0x[i], by = I.py[i] + -1.0 * I.p0y[i];
            sq[0] += bx * bx + by * by;
        }
    } else {
        for (int i = tid; i < n; i += T) sq[0] += I.px[i] * I.px[i] + I.py[i] * I.py[i];
    }
    group_reduce<1>(sq, 0, red);
    if (tid == 0) {
        const double rms = sqrt(sq[0] / (double)(2 * n > 1 ? 2 * n : 1));
        const double dif = sqrt(I.st[0] / (2.0 * I.k));
        const double difid = I.eps * fmax(1e-6, I.facnel * rms);
        const double itgs = I.st[1] + 1.0;
        if (itgs == 1.0) I.st[2] = dif;
        I.st[1] = itgs;
        I.st[0] = 0.0;
        double* r = ro.partials + (size_t)blockIdx.x * 10;
        r[0] = itgs;
        r[1] = dif;
        r[2] = I.st[3];
        r[3] = 0.0;
        r[4] = 0.0;
        r[5] = I.st[2];
        r[6] = (dif >= difid && itgs < (double)I.maxgs) ? 1.0 : 0.0;
        r[7] = difid;
        r[8] = 0.0;
        r[9] = 0.0;
    }
}

// ------------------------------------------------------------------------------------------
// ConvexGS, src/m_solvpt.f90:2446-2821 (CgsItem)
// ------------------------------------------------------------------------------------------

// One wavefront per patch runs every sweep (the element updates are sequential; each shift is
// a direct sum over the interior elements, split over the 64 lanes and reduced by shuffles).
__global__ void __launch_bounds__(64) k_cnvxgs(const CgsItem* items, const int* gstart, int nitems,
                                               RedOut ro)
{
    int it = find_item(gstart, nitems, blockIdx.x);
    const CgsItem& I = items[it];
    const int mx = I.mx, my = I.my, n = I.n, lane = threadIdx.x, M2 = 2 * mx;
    // the interior elements in grid order (iel, src/m_stang.f90:115-124)
    int k = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool in = i < n && I.el[i] >= ADHES;
        const unsigned long long m = __ballot(in);
        const int pos = __popcll(m & ((1ull << lane) - 1ull));
        if (in) I.list[k + pos] = i;
        k += __popcll(m);
    }
    __syncthreads();
    const size_t c0 = (size_t)my * M2 + mx;   // offset (0, 0)
    double coefs[2][2], coefsv[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            coefs[a][b] = I.ga_inv * I.cf[a][b][c0];
            coefsv[a][b] = I.steady ? I.ga_inv * (I.cf[a][b][c0] - I.cv[a][b][c0]) : coefs[a][b];
        }
    // the influence at element i of the current tractions, A(cs) or A(cs - cv) (unscaled)
    auto shift = [&](int i, bool csv, double& ux, double& uy) {
        const int ix = i % mx, iy = i / mx;
        double ax = 0.0, ay = 0.0;
        for (int q = lane; q < k; q += 64) {
            const int j = I.list[q], jx = j % mx, jy = j / mx;
            const size_t c = (size_t)(iy - jy + my) * M2 + (ix - jx + mx);
            double cxx = I.cf[0][0][c], cxy = I.cf[0][1][c], cyx = I.cf[1][0][c], cyy = I.cf[1][1][c];
            if (csv) {
                cxx -= I.cv[0][0][c];
                cxy -= I.cv[0][1][c];
                cyx -= I.cv[1][0][c];
                cyy -= I.cv[1][1][c];
            }
            const double qx = I.px[j], qy = I.py[j];
            ax += cxx * qx + cxy * qy;
            ay += cyx * qx + cyy * qy;
        }
        for (int off = 32; off > 0; off >>= 1) {
            ax += __shfl_xor(ax, off);
            ay += __shfl_xor(ay, off);
        }
        ux = ax;
        uy = ay;
    };
    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0, flops = 0.0;
    while (dif >= difid && itgs < I.maxgs) {
        itgs++;
        double difacc = 0.0;
        if (I.steady) {
            // subnd (src/m_leadedge.f90:336-394) with the edge 1 dx beyond ix0: ubnd = u(ix0 + 1)
            for (int q = 0; q < k; q++) {
                const int i = I.list[q];
                const int j0 = (int)I.lix[i];
                if (j0 < 0) continue;
                double bx = 0.0, by = 0.0;
                if (j0 + 3 <= mx) {
                    shift((i / mx) * mx + j0 + 1, false, bx, by);
                    bx *= I.ga_inv;
                    by *= I.ga_inv;
                }
                if (lane == 0) {
                    I.ubx[i] = bx;
                    I.uby[i] = by;
                }
            }
            __syncthreads();
        }
        for (int q = 0; q < k; q++) {
            const int i = I.list[q];
            const int8_t e0 = I.el[i];
            // after 1000 sweeps the slip elements are skipped on odd sweeps (:2603)
            if (!(itgs <= 1000 || itgs % 2 == 0 || e0 == ADHES)) continue;
            const bool zledge = I.steady && I.lix[i] >= 0.0;
            double ux, uy;
            shift(i, I.steady && !zledge, ux, uy);
            double sv[2] = {I.wsx[i] + I.ga_inv * ux, I.wsy[i] + I.ga_inv * uy};
            if (zledge) {
                sv[0] -= I.ubx[i];
                sv[1] -= I.uby[i];
            }
            const double (*cc)[2] = (!I.steady || zledge) ? coefs : coefsv;
            const double opx = I.px[i], opy = I.py[i];
            double pr[3] = {opx, opy, I.pn[i]};
            int8_t e8 = e0;
            elmtrc(e8, cc[0][0], cc[0][1], cc[1][0], cc[1][1], I.eps, pr, I.mu, sv, I.omegah, I.omegas, false);
            difacc += (opx - pr[0]) * (opx - pr[0]) + (opy - pr[1]) * (opy - pr[1]);
            if (lane == 0) {
                I.px[i] = pr[0];
                I.py[i] = pr[1];
                I.sx[i] = sv[0];
                I.sy[i] = sv[1];
                I.el[i] = e8;
            }
            __syncthreads();
        }
        flops += 8.0 * (double)k * (double)k;
        double sq = 0.0;
        for (int i = lane; i < n; i += 64) {
            if (I.p0x) {
                const double bx = I.px[i] + -1.0 * I.p0x[i], by = I.py[i] + -1.0 * I.p0y[i];
                sq += bx * bx + by * by;
            } else {
                sq += I.px[i] * I.px[i] + I.py[i] * I.py[i];
            }
        }
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
        const double rms = sqrt(sq / (double)(2 * n > 1 ? 2 * n : 1));
        dif = sqrt(difacc / (2.0 * I.k));
        difid = I.eps * fmax(1e-6, I.facnel * rms);
        if (itgs == 1) dif1 = dif;
    }
    if (lane == 0) {
        double* r = ro.partials + (size_t)blockIdx.x * 10;
        r[0] = (double)itgs;
        r[1] = dif;
        r[2] = flops;
        r[3] = 0.0;
        r[4] = 0.0;
        r[5] = dif1;
        for (int q = 6; q < 10; q++) r[q] = 0.0;
    }
}

// ------------------------------------------------------------------------------------------
// Dense influence product on FP64 MFMA (DenseItem)
// ------------------------------------------------------------------------------------------

typedef double dense_v4 __attribute__((ext_vector_type(4)));

// One work-group per tile of 16 x 16 targets (ix0 + i, iy0 + j).  With d = iy - jy fixed the
// product is an ordinary matrix product: the Toeplitz block T_d (T_d[ix][jx] = cf(ix - jx, d))
// times the source rows jy = iy - d.  v_mfma_f64_16x16x4f64 takes A = 16 targets x 4 sources
// (lane i + 16 k holds A[i][k]) and B = 4 sources x 16 target rows (lane j + 16 k holds B[k][j]);
// register r of lane j + 16 q holds D[4 r + q][j], r = 0..3 (measured on gfx950).  The k loop
// runs over d and the sources in steps of 4; coefficients outside the tables and sources
// outside the grid enter as zeros.  Each of the group's wavefronts takes a contiguous part of
// the d range (a serial chain an eighth as long); wavefront 0 adds the parts in wave order.
constexpr int DENSE_WAVES = 8;

__global__ void __launch_bounds__(64 * DENSE_WAVES) k_dense(const DenseItem* items, const int* gstart, int nitems)
{
    __shared__ double part[DENSE_WAVES][3][4][64];
    const int it = find_item(gstart, nitems, blockIdx.x);
    const DenseItem& I = items[it];
    const int mx = I.mx, my = I.my, M2 = 2 * mx, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntx = (mx + 15) / 16;
    const int t = blockIdx.x - gstart[it];
    const int ix0 = (t % ntx) * 16, iy0 = (t / ntx) * 16;
    const int i = lane & 15, q = lane >> 4;
    const int nin = I.nin, nout = I.nout;
    // one accumulator per (out, in) block: independent MFMA chains, summed at the end
    dense_v4 acc[3][3];
#pragma unroll
    for (int o = 0; o < 3; o++)
#pragma unroll
        for (int j = 0; j < 3; j++) acc[o][j] = (dense_v4){0.0, 0.0, 0.0, 0.0};
    const int ia = ix0 + i;       // this lane's target in A (row i)
    const int jb = iy0 + i;       // this lane's target row in B (column i)
    // the tile's row offsets, split in DENSE_WAVES contiguous parts
    const int dlo = max(iy0 - (my - 1), -(my - 1)), dhi = min(iy0 + 15, my - 1);
    const int nd = dhi - dlo + 1, per = (nd + DENSE_WAVES - 1) / DENSE_WAVES;
    const int d0 = dlo + wave * per, d1 = min(dhi, d0 + per - 1);
    const int nstep = (mx + 3) / 4;
    // the operands of step (d, jx0) for this lane: A of every block, B of every input
    auto fetch = [&](int d, int jx0, double (&a)[3][3], double (&b)[3]) {
        const int jy = jb - d;
        const int jx = jx0 + q;
        const bool aok = jx < mx && ia < mx;
        const bool bok = jx < mx && jb < my && jy >= 0 && jy < my;
        const size_t ca = aok ? (size_t)(d + my) * M2 + (size_t)(ia - jx + mx) : 0;
        const size_t sb = bok ? (size_t)jy * mx + jx : 0;
#pragma unroll
        for (int j = 0; j < 3; j++) b[j] = (j < nin && bok) ? I.in[j][sb] : 0.0;
#pragma unroll
        for (int o = 0; o < 3; o++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const double* c = (o < nout && j < nin) ? I.cf[o][j] : nullptr;
                a[o][j] = (c && aok) ? c[ca] : 0.0;
            }
    };
    if (d0 <= d1) {
        // software pipeline: the loads of the next step are issued before this step's MFMAs
        double a[3][3], b[3], an[3][3], bn[3];
        int d = d0, k = 0;
        fetch(d, 0, a, b);
        for (;;) {
            int dn = d, kn = k + 1;
            if (kn == nstep) {
                kn = 0;
                dn = d + 1;
            }
            const bool more = dn <= d1;
            if (more) fetch(dn, 4 * kn, an, bn);
#pragma unroll
            for (int o = 0; o < 3; o++) {
                if (o >= nout) break;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    if (j >= nin || !I.cf[o][j]) continue;
                    acc[o][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[o][j], b[j], acc[o][j], 0, 0, 0);
                }
            }
            if (!more) break;
            d = dn;
            k = kn;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                b[j] = bn[j];
#pragma unroll
                for (int o = 0; o < 3; o++) a[o][j] = an[o][j];
            }
        }
    }
    // each wavefront's sums over its blocks, then wavefront 0 adds the parts in wave order
#pragma unroll
    for (int o = 0; o < 3; o++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double u = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (j < nin) u += acc[o][j][r];
            part[wave][o][r][lane] = u;
        }
    __syncthreads();
    if (wave != 0) return;
    const int iy = iy0 + i;
    if (iy >= my) return;
#pragma unroll
    for (int o = 0; o < 3; o++) {
        if (o >= nout) break;
        double* dst = I.out[o];
        const double* add = I.addend[o];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ix = ix0 + 4 * r + q;
            if (ix >= mx) continue;
            const size_t ii = (size_t)iy * mx + ix;
            if (I.mask != M_ALL && !in_mask(I.mask, I.el[ii])) continue;
            double u = 0.0;
#pragma unroll
            for (int w = 0; w < DENSE_WAVES; w++) u += part[w][o][r][lane];
            u *= I.ga_inv;
            if (add) u += add[ii];
            dst[ii] = u;
        }
    }
}

// ------------------------------------------------------------------------------------------
// host-side launchers
// ------------------------------------------------------------------------------------------

static int ngroups_elem(int n) { return (n + ELEM_THREADS * ELEM_PER_THREAD - 1) / (ELEM_THREADS * ELEM_PER_THREAD); }
static int ngroups_rows(int my) { return (my + ROW_THREADS - 1) / ROW_THREADS; }

int op_nred(int kind, const void* item)
{
    switch (kind) {
    case OP_VEC: return 4;
    case OP_NCG_NEGPN: return 3;
    case OP_NCG_DD: return 2;
    case OP_NCG_EXPAND: return 1;
    case OP_NCG_HMIN: return 1;
    case OP_SN_CONTRACT: return 2;
    case OP_SN_EXPAND: return 2;
    case OP_SN_FINAL: return 1;
    case OP_AIJ_ONES: return 1;
    case OP_TCG_ZPREP: return 2;
    case OP_TCG_QPOST: return 2;
    case OP_TCG_PUPD: return 4;
    case OP_TCG_AS: return 1;
    case OP_TCG_SA: return 2;
    case OP_ST_AS: return 1;
    case OP_ST_SA: return 1;
    case OP_GD_PROBE: return 10;
    case OP_GSB_SETUP: case OP_GSB_CHECK: return 10;
    case OP_CNVXGS: return 10;
    case OP_GD_MOVES: return 1;
    case OP_GD_DIAGRES: return 3;
    case OP_SETUP: return 2;
    case OP_FORCES: return 5;
    case OP_STDYGS: return 10;
    case OP_GUARD: return 5;
    case OP_SENS_EXT: return 4;
    default: return 0;
    }
    (void)item;
}

int op_combine_mode(int kind, const void* item)
{
    if (kind == OP_VEC) {
        int sub = ((const VecItem*)item)->sub;
        if (sub == V_MAXABS || sub == V_MAX) return 1;
        if (sub == V_MIN) return 2;
    }
    if (kind == OP_SETUP || kind == OP_SENS_EXT) return 1;
    return 0;
}

int op_groups(int kind, const void* item, int conv_stage)
{
    switch (kind) {
    case OP_VEC: return ngroups_elem(((const VecItem*)item)->n);
    case OP_NCG_NEGPN: case OP_NCG_DD: case OP_NCG_EXPAND: case OP_NCG_HMIN:
    case OP_SN_CONTRACT: case OP_SN_EXPAND: case OP_SN_FINAL:
        return ngroups_elem(((const NcgItem*)item)->n);
    case OP_AIJ_ONES: return ngroups_rows(((const AijItem*)item)->my);
    case OP_TCG_NT: case OP_TCG_RSET: case OP_TCG_ZPREP: case OP_TCG_QPOST: case OP_TCG_PUPD:
    case OP_TCG_AS: case OP_TCG_SA: case OP_TCG_VUPD:
        return ngroups_elem(((const TcgItem*)item)->n);
    case OP_ST_AS: case OP_ST_SA: return ngroups_elem(((const StItem*)item)->n);
    case OP_GD_DP: return ngroups_elem(((const GdItem*)item)->n);
    case OP_GD_PROJECT: case OP_GD_PROBE: case OP_GD_UPDATE: {
        const GdItem* g = (const GdItem*)item;
        int R = gd_rows_per_tile(g->mx, g->my, kind == OP_GD_PROJECT ? GD_TILE_PROJ : GD_TILE);
        return (g->my + R - 1) / R;
    }
    case OP_GD_MOVES: case OP_GD_DIAGRES: return ngroups_elem(((const GdItem*)item)->n);
    case OP_SETUP: case OP_FORCES: return ngroups_elem(((const SetupItem*)item)->n);
    case OP_STDYGS: return 1;
    case OP_GSB_SETUP: case OP_GSB_WALK: case OP_GSB_CHECK: return 1;
    case OP_CNVXGS: return 1;
    case OP_DENSE: {
        const DenseItem* d = (const DenseItem*)item;
        return ((d->mx + 15) / 16) * ((d->my + 15) / 16);
    }
    case OP_SENS: case OP_SENS_EXT: return ngroups_elem(((const SensItem*)item)->n);
    case OP_GUARD: return (((const GsItem*)item)->my + ELEM_THREADS - 1) / ELEM_THREADS;
    case OP_LEDGE: {
        const LedgeItem* l = (const LedgeItem*)item;
        return l->mode == 0 ? (l->my + ELEM_THREADS - 1) / ELEM_THREADS : ngroups_elem(l->n);
    }
    case OP_ELASCF: { const CoefItem* c = (const CoefItem*)item; return (int)(((size_t)(2 * c->mx) * (2 * c->my) + ELEM_THREADS - 1) / ELEM_THREADS); }
    case OP_EMBED: { const EmbedItem* e = (const EmbedItem*)item; return (int)(((size_t)e->Lx * e->Ly + ELEM_THREADS - 1) / ELEM_THREADS); }
    case OP_EXTRACT: { const ExtractItem* e = (const ExtractItem*)item; return (int)(((size_t)(2 * e->cf_mx) * (2 * e->cf_my) + ELEM_THREADS - 1) / ELEM_THREADS); }
    case OP_CONV: case OP_CONV32: {
        const ConvItem* c = (const ConvItem*)item;
        int M = c->Lx / 2, H = M + 1;
        int nrow = c->mode == 0 ? c->my : c->Ly;
        int nout = c->mode == 0 ? c->nout : c->nin;
        const int rpg = conv_rows(*c);
        (void)M;
        if (conv_stage == 0) return c->nin * ((nrow + rpg - 1) / rpg);
        if (conv_stage == 1) {
            int nc = c->nin > nout ? c->nin : nout;
            int cpg = conv_cols(*c, nc);
            return (H + cpg - 1) / cpg;
        }
        if (c->mode == 1) return 0;
        return nout * ((nrow + rpg - 1) / rpg);
    }
    }
    return 0;
}

static size_t conv_lds(int kind, const ConvItem* c, int stage)
{
    size_t es = kind == OP_CONV32 ? 8 : 16;
    int M = c->Lx / 2;
    if (stage == 0 || stage == 2) return 2 * (size_t)conv_rows(*c) * M * es;
    int nout = c->mode == 0 ? c->nout : c->nin;
    int nc = c->nin > nout ? c->nin : nout;
    return 2 * (size_t)nc * conv_cols(*c, nc) * c->Ly * es;
}

void launch_op(int kind, const void* d_items, const int* d_gstart, int nitems, int ngroups,
               RedOut ro, int conv_stage, size_t lds_bytes, hipStream_t s)
{
    if (ngroups <= 0) return;
    dim3 ge(ngroups), te(ELEM_THREADS), tr(ROW_THREADS), tf(FFT_THREADS);
    switch (kind) {
    case OP_VEC: hipLaunchKernelGGL(k_vec, ge, te, 0, s, (const VecItem*)d_items, d_gstart, nitems, ro); break;
    case OP_NCG_NEGPN: hipLaunchKernelGGL(k_ncg_negpn, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_NCG_DD: hipLaunchKernelGGL(k_ncg_dd, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_NCG_EXPAND: hipLaunchKernelGGL(k_ncg_expand, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_NCG_HMIN: hipLaunchKernelGGL(k_ncg_hmin, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_SN_CONTRACT: hipLaunchKernelGGL(k_sn_contract, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_SN_EXPAND: hipLaunchKernelGGL(k_sn_expand, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_SN_FINAL: hipLaunchKernelGGL(k_sn_final, ge, te, 0, s, (const NcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_AIJ_ONES: hipLaunchKernelGGL(k_aij_ones, ge, tr, 0, s, (const AijItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_NT: hipLaunchKernelGGL(k_tcg_nt, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_RSET: hipLaunchKernelGGL(k_tcg_rset, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_ZPREP: hipLaunchKernelGGL(k_tcg_zprep, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_QPOST: hipLaunchKernelGGL(k_tcg_qpost, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_PUPD: hipLaunchKernelGGL(k_tcg_pupd, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_VUPD: hipLaunchKernelGGL(k_tcg_vupd, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_AS: hipLaunchKernelGGL(k_tcg_as, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_TCG_SA: hipLaunchKernelGGL(k_tcg_sa, ge, te, 0, s, (const TcgItem*)d_items, d_gstart, nitems, ro); break;
    case OP_ST_AS: hipLaunchKernelGGL(k_st_as, ge, te, 0, s, (const StItem*)d_items, d_gstart, nitems, ro); break;
    case OP_ST_SA: hipLaunchKernelGGL(k_st_sa, ge, te, 0, s, (const StItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_DP: hipLaunchKernelGGL(k_gd_dp, ge, te, 0, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_CNVXGS: hipLaunchKernelGGL(k_cnvxgs, ge, dim3(64), 0, s, (const CgsItem*)d_items, d_gstart, nitems, ro); break;
    case OP_DENSE: hipLaunchKernelGGL(k_dense, ge, dim3(64 * DENSE_WAVES), 0, s, (const DenseItem*)d_items, d_gstart, nitems); break;
    case OP_GSB_SETUP: hipLaunchKernelGGL(k_gsb_setup, ge, dim3(GSB_THREADS), 0, s, (const GsbItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GSB_WALK: hipLaunchKernelGGL(k_gsb_walk, ge, dim3(GSB_THREADS), lds_bytes, s, (const GsbItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GSB_CHECK: hipLaunchKernelGGL(k_gsb_check, ge, dim3(GSB_THREADS), 0, s, (const GsbItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_PROJECT: hipLaunchKernelGGL(k_gd_project, ge, dim3(GD_THREADS), lds_bytes, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_PROBE: hipLaunchKernelGGL(k_gd_probe, ge, dim3(GD_THREADS), lds_bytes, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_UPDATE: hipLaunchKernelGGL(k_gd_update, ge, dim3(GD_THREADS), lds_bytes, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_MOVES: hipLaunchKernelGGL(k_gd_moves, ge, te, 0, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_GD_DIAGRES: hipLaunchKernelGGL(k_gd_diagres, ge, te, 0, s, (const GdItem*)d_items, d_gstart, nitems, ro); break;
    case OP_SETUP: hipLaunchKernelGGL(k_setup, ge, te, 0, s, (const SetupItem*)d_items, d_gstart, nitems, ro); break;
    case OP_FORCES: hipLaunchKernelGGL(k_forces, ge, te, 0, s, (const SetupItem*)d_items, d_gstart, nitems, ro); break;
    case OP_STDYGS:
        if (conv_stage == 2) hipLaunchKernelGGL(k_stdygs3, ge, dim3(GS_THREADS), lds_bytes, s, (const GsItem*)d_items, d_gstart, nitems, ro);
        else if (conv_stage == 1) hipLaunchKernelGGL(k_stdygs2, ge, dim3(GS_THREADS), lds_bytes, s, (const GsItem*)d_items, d_gstart, nitems, ro);
        else hipLaunchKernelGGL(k_stdygs, ge, dim3(GS_THREADS), lds_bytes, s, (const GsItem*)d_items, d_gstart, nitems, ro);
        break;
    case OP_GUARD: hipLaunchKernelGGL(k_guard, ge, te, 0, s, (const GsItem*)d_items, d_gstart, nitems, ro); break;
    case OP_SENS: case OP_SENS_EXT: hipLaunchKernelGGL(k_sens, ge, te, 0, s, (const SensItem*)d_items, d_gstart, nitems, ro); break;
    case OP_LEDGE: hipLaunchKernelGGL(k_ledge, ge, te, 0, s, (const LedgeItem*)d_items, d_gstart, nitems, ro); break;
    case OP_ELASCF: hipLaunchKernelGGL(k_elascf, ge, te, 0, s, (const CoefItem*)d_items, d_gstart, nitems); break;
    case OP_EMBED: hipLaunchKernelGGL(k_embed, ge, te, 0, s, (const EmbedItem*)d_items, d_gstart, nitems); break;
    case OP_EXTRACT: hipLaunchKernelGGL(k_extract, ge, te, 0, s, (const ExtractItem*)d_items, d_gstart, nitems); break;
    case OP_CONV:
        if (conv_stage == 0) hipLaunchKernelGGL(k_rowfwd<double>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        else if (conv_stage == 1) hipLaunchKernelGGL(k_colmul<double>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        else hipLaunchKernelGGL(k_rowinv<double>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        break;
    case OP_CONV32:
        if (conv_stage == 0) hipLaunchKernelGGL(k_rowfwd<float>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        else if (conv_stage == 1) hipLaunchKernelGGL(k_colmul<float>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        else hipLaunchKernelGGL(k_rowinv<float>, ge, tf, lds_bytes, s, (const ConvItem*)d_items, d_gstart, nitems);
        break;
    }
}

// Algorithmic HBM bytes of one item of `kind` in `stage`: every array the kernel must read or
// write, each element once (FFT intermediates of the three product stages included as the
// stages' inputs/outputs).  Used for the achieved-bandwidth report only.
double op_bytes(int kind, const void* item, int stage)
{
    switch (kind) {
    case OP_VEC: {
        const VecItem* v = (const VecItem*)item;
        static const int arrays[V_NSUB] = {1, 2, 2, 3, 3, 2, 2, 1, 1, 1, 1, 1, 0};
        return (double)v->n * (1.0 + 8.0 * v->ncomp * arrays[v->sub]);
    }
    case OP_NCG_NEGPN: case OP_NCG_HMIN: case OP_SN_CONTRACT: case OP_SN_FINAL:
        return ((const NcgItem*)item)->n * 18.0;
    case OP_NCG_DD: return ((const NcgItem*)item)->n * 25.0;
    case OP_NCG_EXPAND: return ((const NcgItem*)item)->n * 34.0;
    case OP_SN_EXPAND: return ((const NcgItem*)item)->n * 17.0;
    case OP_TCG_NT: return ((const TcgItem*)item)->n * 32.0;
    case OP_TCG_RSET: return ((const TcgItem*)item)->n * 49.0;
    case OP_TCG_ZPREP: return ((const TcgItem*)item)->n * 105.0;
    case OP_TCG_QPOST: return ((const TcgItem*)item)->n * 121.0;
    case OP_TCG_PUPD: return ((const TcgItem*)item)->n * 89.0;
    case OP_TCG_AS: return ((const TcgItem*)item)->n * 26.0;
    case OP_TCG_SA: return ((const TcgItem*)item)->n * 82.0;
    case OP_ST_AS: case OP_ST_SA: return ((const StItem*)item)->n * 42.0;
    case OP_GD_DP: { const GdItem* g = (const GdItem*)item; return (double)g->n * 32.0; }
    case OP_GD_PROJECT: { const GdItem* g = (const GdItem*)item; return (double)g->n * 73.0; }
    // six 2N FP64 vectors per probe (the algorithmic count: dp, dv, s, q, n, v), two more (r)
    // on the first probe of a line search
    case OP_GD_PROBE: { const GdItem* g = (const GdItem*)item; return (double)g->n * (96.0 + (g->first ? 32.0 : 0.0)); }
    case OP_GD_UPDATE: { const GdItem* g = (const GdItem*)item; return (double)g->n * 121.0; }
    case OP_GD_MOVES: { const GdItem* g = (const GdItem*)item; return (double)g->n * 66.0; }
    case OP_GD_DIAGRES: { const GdItem* g = (const GdItem*)item; return (double)g->n * 121.0; }
    case OP_SETUP: { const SetupItem* s = (const SetupItem*)item; return (double)s->n * 97.0; }
    case OP_FORCES: { const SetupItem* s = (const SetupItem*)item; return (double)s->n * 24.0; }
    case OP_CONV: case OP_CONV32: {
        const ConvItem* c = (const ConvItem*)item;
        double es = kind == OP_CONV32 ? 8.0 : 16.0;
        int M = c->Lx / 2, H = M + 1;
        int nrow = c->mode == 0 ? c->my : c->Ly;
        int nxr = c->mode == 0 ? c->mx : c->Lx;
        int nout = c->mode == 0 ? c->nout : c->nin;
        // compulsory traffic of one product group: the inputs (8 B per point, row stage), the
        // coefficient spectra (column stage) and the outputs (+ addend, + states; inverse row
        // stage).  The half-spectrum round trips between the stages are not counted.
        double npt = (double)nrow * nxr;
        (void)nout;
        if (stage == 0) return c->nin * npt * 8.0;
        if (stage == 1) {
            if (c->mode == 1) return c->nin * (double)H * c->Ly * es;   // spectrum store
            if (c->mode == 2) return 0.0;
            int nspec = 0;
            for (int i = 0; i < c->nout; i++)
                for (int j = 0; j < c->nin; j++) nspec += c->spec[i][j] != nullptr;
            return nspec * (double)H * c->Ly * es;
        }
        if (c->mode == 1) return 0.0;
        double outb = 8.0 + (c->mode == 0 ? 1.0 : 0.0);
        for (int i = 0; i < c->nout && c->mode == 0; i++)
            if (c->addend[i]) outb += 8.0 / c->nout;
        return (c->mode == 0 ? c->nout : c->nin) * npt * outb;
    }
    case OP_ELASCF: { const CoefItem* c = (const CoefItem*)item; return 9.0 * 8.0 * (2 * c->mx) * (2 * c->my); }
    case OP_DENSE: {
        // FP64 flops of the dense matrix-vector products: 2 n^2 per nonzero block
        const DenseItem* d = (const DenseItem*)item;
        int nblk = 0;
        for (int o = 0; o < d->nout; o++)
            for (int j = 0; j < d->nin; j++) nblk += d->cf[o][j] != nullptr;
        const double n = (double)d->mx * d->my;
        return 2.0 * n * n * nblk;
    }
    case OP_STDYGS: { const GsItem* g = (const GsItem*)item; return (double)g->n * 73.0; }
    case OP_GUARD: { const GsItem* g = (const GsItem*)item; return (double)g->my; }
    case OP_SENS: {
        const SensItem* q = (const SensItem*)item;
        if (q->mode == 0) return (double)q->n * ((q->el_dst ? 2.0 : 0.0) + 16.0 * q->nvec);
        return (double)q->n * 33.0;
    }
    case OP_SENS_EXT: return (double)((const SensItem*)item)->n;
    case OP_EMBED: { const EmbedItem* e = (const EmbedItem*)item; return 8.0 * e->Lx * e->Ly; }
    case OP_EXTRACT: { const ExtractItem* e = (const ExtractItem*)item; return 16.0 * (2 * e->cf_mx) * (2 * e->cf_my); }
    }
    return 0.0;
}

size_t op_lds(int kind, const void* item, int stage)
{
    if (kind == OP_CONV || kind == OP_CONV32) return conv_lds(kind, (const ConvItem*)item, stage);
    if (kind == OP_GSB_WALK) return gsb_lds(((const GsbItem*)item)->mx);
    if (kind == OP_GD_PROJECT || kind == OP_GD_PROBE || kind == OP_GD_UPDATE) {
        const GdItem* g = (const GdItem*)item;
        int R = gd_rows_per_tile(g->mx, g->my, kind == OP_GD_PROJECT ? GD_TILE_PROJ : GD_TILE);
        size_t cells = (size_t)R * gd_row_stride(g->mx);
        return cells * ((kind == OP_GD_PROJECT ? 7 : 3) * sizeof(double) + 1) + 16;
    }
    if (kind == OP_STDYGS) {
        const GsItem* g = (const GsItem*)item;
        if (stage == 3) return g->mx >= 2 ? gsb_lds(g->mx) : MAX_LDS + 1;   // rng needs 4 my <= 2 n
        if (stage == 2) return gs3_lds(g->mx, g->my);
        if (stage == 1) return gs2_lds(g->mx, g->my);
        return (size_t)g->mx * (5 * sizeof(double) + 1) + (size_t)(2 * g->my + 1) * sizeof(int) + 16;
    }
    return 0;
}

__global__ void k_fill_u32(unsigned* p, size_t n, unsigned v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s)
{
    if (n == 0) return;
    size_t g = (n + 255) / 256;
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
}

void conv_default_shape(const ConvItem* c, int* rpg, int* cpg)
{
    const int nout = c->mode == 0 ? c->nout : c->nin;
    const int nc = c->nin > nout ? c->nin : nout;
    *rpg = rows_per_group(c->Lx / 2);
    *cpg = cols_per_group(c->Ly, nc);
}

void launch_rollcoef(int n, const double* cs, double* cv, hipStream_t s)
{
    int g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_rollcoef, dim3(g), dim3(256), 0, s, n, cs, cv);
}



void launch_to_float(size_t n, const double* in, float* out, hipStream_t s)
{
    int g = (int)((n + 255) / 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_to_float, dim3(g), dim3(256), 0, s, n, in, out);
}

}  // namespace cntc
