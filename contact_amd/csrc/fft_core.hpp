// fft_core.hpp -- mixed-radix Stockham FFT passes for transforms staged in LDS.
//
// A work-group holds `nfft` transforms of length L side by side (transform t at buf + t*L)
// and runs one pass per radix.  Every pass reads one buffer and writes the other (autosort),
// so no digit reversal is needed.  Radices 2, 3, 4, 5, 7 and 8 have register butterflies; any
// other prime factor runs a generic pass (lengths of the preconditioner's 2*mx x 2*my circulant).  The functions are __host__ __device__ so the
// index arithmetic can be checked on the CPU against numpy with tid=0, nthreads=1.
#pragma once

#include <hip/hip_runtime.h>

namespace cntc {

template <typename R>
struct cpx {
    R x, y;
};

template <typename R>
__host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__host__ __device__ inline cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__host__ __device__ inline cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__host__ __device__ inline cpx<R> cconj(cpx<R> a) { return {a.x, -a.y}; }
// multiply by -i*dir*(-1) helpers: rot(a, s) = a * (s*i) for s = +1/-1
template <typename R>
__host__ __device__ inline cpx<R> mul_i(cpx<R> a, R s) { return {-s * a.y, s * a.x}; }

// Plan of one length: radices in the order they are applied.
struct FftPlan {
    int L;
    int npass;
    int rad[16];
};

__host__ __device__ inline FftPlan make_plan(int L)
{
    FftPlan p;
    p.L = L;
    p.npass = 0;
    int n = L;
    while (n % 8 == 0 && p.npass < 16) { p.rad[p.npass++] = 8; n /= 8; }
    while (n % 4 == 0 && p.npass < 16) { p.rad[p.npass++] = 4; n /= 4; }
    while (n % 2 == 0 && p.npass < 16) { p.rad[p.npass++] = 2; n /= 2; }
    while (n % 3 == 0 && p.npass < 16) { p.rad[p.npass++] = 3; n /= 3; }
    while (n % 5 == 0 && p.npass < 16) { p.rad[p.npass++] = 5; n /= 5; }
    while (n % 7 == 0 && p.npass < 16) { p.rad[p.npass++] = 7; n /= 7; }
    // remaining prime factors: generic passes (O(L*R) work, used for preconditioner sizes)
    for (int f = 11; n > 1 && p.npass < 16;) {
        if ((long)f * f > n) f = n;
        if (n % f == 0) { p.rad[p.npass++] = f; n /= f; }
        else f += 2;
    }
    if (n != 1) p.npass = -1;
    return p;
}

// Small DFT of radix R in registers; sgn = -1 forward, +1 inverse.
template <typename R, int RAD>
__host__ __device__ inline void small_dft(cpx<R>* v, R sgn)
{
    if constexpr (RAD == 2) {
        cpx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (RAD == 4) {
        cpx<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cpx<R> b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        cpx<R> b1r = mul_i(b1, sgn);   // sgn*i * b1
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1r);
        v[3] = csub(a1, b1r);
    } else if constexpr (RAD == 8) {
        const R h = (R)0.70710678118654752440084436210485;
        // radix-2 over stride 4, then twiddles, then two radix-4
        cpx<R> a[4], b[4];
        for (int r = 0; r < 4; r++) {
            a[r] = cadd(v[r], v[r + 4]);
            b[r] = csub(v[r], v[r + 4]);
        }
        // b[r] *= exp(sgn * 2 pi i r / 8)
        b[1] = {h * (b[1].x - sgn * b[1].y), h * (b[1].y + sgn * b[1].x)};
        b[2] = mul_i(b[2], sgn);
        b[3] = {h * (-b[3].x - sgn * b[3].y), h * (-b[3].y + sgn * b[3].x)};
        small_dft<R, 4>(a, sgn);
        small_dft<R, 4>(b, sgn);
        for (int r = 0; r < 4; r++) {
            v[2 * r] = a[r];
            v[2 * r + 1] = b[r];
        }
    } else {
        // radix 3, 5, 7: direct DFT with the roots exp(2 pi i e / RAD) as constants
        constexpr double c3[3] = {1.00000000000000000000e+00, -4.99999999999999777955e-01, -5.00000000000000444089e-01};
        constexpr double s3[3] = {0.00000000000000000000e+00, 8.66025403784438707611e-01, -8.66025403784438374544e-01};
        constexpr double c5[5] = {1.00000000000000000000e+00, 3.09016994374947451263e-01, -8.09016994374947340241e-01, -8.09016994374947562285e-01, 3.09016994374947229218e-01};
        constexpr double s5[5] = {0.00000000000000000000e+00, 9.51056516295153531182e-01, 5.87785252292473248126e-01, -5.87785252292473026081e-01, -9.51056516295153642204e-01};
        constexpr double c7[7] = {1.00000000000000000000e+00, 6.23489801858733594386e-01, -2.22520933956314337365e-01, -9.00968867902419034976e-01, -9.00968867902419145999e-01, -2.22520933956314587165e-01, 6.23489801858733372342e-01};
        constexpr double s7[7] = {0.00000000000000000000e+00, 7.81831482468029803634e-01, 9.74927912181823619342e-01, 4.33883739117558231424e-01, -4.33883739117558009379e-01, -9.74927912181823619342e-01, -7.81831482468029914656e-01};
        const double* cr = RAD == 3 ? c3 : (RAD == 5 ? c5 : c7);
        const double* sr = RAD == 3 ? s3 : (RAD == 5 ? s5 : s7);
        cpx<R> out[RAD];
        for (int k = 0; k < RAD; k++) {
            cpx<R> acc = v[0];
            for (int r = 1; r < RAD; r++) {
                int e = (r * k) % RAD;
                cpx<R> w = {(R)cr[e], (R)(sgn * (R)sr[e])};
                acc = cadd(acc, cmul(v[r], w));
            }
            out[k] = acc;
        }
        for (int k = 0; k < RAD; k++) v[k] = out[k];
    }
}

// One Stockham pass of radix RAD over nfft transforms of length L.
// tw[e] = exp(-2 pi i e / L); the inverse direction uses the conjugates.
template <typename R, int RAD>
__host__ __device__ inline void fft_pass(const cpx<R>* src, cpx<R>* dst, int L, int nfft, int Ns,
                                         R sgn, const cpx<R>* tw, int tid, int nthreads)
{
    const int nb = L / RAD;
    const int stride_e = L / (Ns * RAD);
    const int total = nb * nfft;
    for (int w = tid; w < total; w += nthreads) {
        int t = w / nb;
        int j = w - t * nb;
        int k = j % Ns;
        const cpx<R>* s = src + (size_t)t * L;
        cpx<R>* d = dst + (size_t)t * L;
        cpx<R> v[RAD];
        v[0] = s[j];
        for (int r = 1; r < RAD; r++) {
            cpx<R> a = s[j + r * nb];
            cpx<R> wv = tw[r * k * stride_e];
            if (sgn > 0) wv.y = -wv.y;
            v[r] = cmul(a, wv);
        }
        small_dft<R, RAD>(v, sgn);
        int base = (j / Ns) * Ns * RAD + k;
        for (int r = 0; r < RAD; r++) d[base + r * Ns] = v[r];
    }
}

// Generic Stockham pass for any radix RAD (each work item forms one output as an RAD-term sum).
template <typename R>
__host__ __device__ inline void fft_pass_generic(const cpx<R>* src, cpx<R>* dst, int L, int nfft,
                                                 int Ns, int RAD, R sgn, const cpx<R>* tw, int tid,
                                                 int nthreads)
{
    const int nb = L / RAD;
    const int stride_e = L / (Ns * RAD);
    const int total = L * nfft;
    for (int w = tid; w < total; w += nthreads) {
        int t = w / L;
        int rem = w - t * L;
        int j = rem / RAD;
        int r_out = rem - j * RAD;
        int k = j % Ns;
        const cpx<R>* s = src + (size_t)t * L;
        int step = (int)(((long)k * stride_e + (long)r_out * nb) % L);
        int e = 0;
        cpx<R> acc = {(R)0, (R)0};
        for (int r = 0; r < RAD; r++) {
            cpx<R> wv = tw[e];
            if (sgn > 0) wv.y = -wv.y;
            acc = cadd(acc, cmul(s[j + r * nb], wv));
            e += step;
            if (e >= L) e -= L;
        }
        int base = (j / Ns) * Ns * RAD + k;
        dst[(size_t)t * L + base + r_out * Ns] = acc;
    }
}

// Full transform; returns the buffer that holds the result (a or b).  Between passes the
// caller's barrier `sync()` is invoked.
template <typename R, typename Sync>
__host__ __device__ inline cpx<R>* fft_run(cpx<R>* a, cpx<R>* b, const FftPlan& plan, int nfft,
                                           R sgn, const cpx<R>* tw, int tid, int nthreads,
                                           Sync sync)
{
    int Ns = 1;
    for (int ps = 0; ps < plan.npass; ps++) {
        int rad = plan.rad[ps];
        switch (rad) {
        case 2: fft_pass<R, 2>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        case 3: fft_pass<R, 3>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        case 4: fft_pass<R, 4>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        case 5: fft_pass<R, 5>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        case 7: fft_pass<R, 7>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        case 8: fft_pass<R, 8>(a, b, plan.L, nfft, Ns, sgn, tw, tid, nthreads); break;
        default: fft_pass_generic<R>(a, b, plan.L, nfft, Ns, rad, sgn, tw, tid, nthreads); break;
        }
        sync();
        cpx<R>* t = a;
        a = b;
        b = t;
        Ns *= rad;
    }
    return a;
}

}  // namespace cntc
