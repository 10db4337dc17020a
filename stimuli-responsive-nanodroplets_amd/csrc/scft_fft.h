// scft_fft.h — LDS-resident Stockham FFT passes for gfx950 (device code).
//
// A workgroup owns P complex lines of length L in LDS (two real grid lines
// packed per complex line: the propagator's operator is real, so the real
// and imaginary parts advance independently).  Passes run Stockham autosort
// (natural-order output, no bit reversal) with ping-pong buffers:
//   for each radix R of the plan, with Ns = product of earlier radices:
//     for j in [0, L/R):  k = j mod Ns
//       v[r] = x[j + r L/R] * w_L^{sign r k L/(Ns R)},  v = DFT_R(v)
//       y[(j/Ns) Ns R + k + r Ns] = v[r]
// Radices 2, 3, 4 and 5 use register butterflies; any other factor (7, 13,
// 43, ...) uses a direct DFT pass in which each thread forms one output.
#pragma once

#include <hip/hip_runtime.h>

namespace scft {

constexpr int kMaxPass = 12;

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }

template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }

// s * i * a for s = +-1
template <typename T>
__device__ __forceinline__ cpx<T> cmul_si(cpx<T> a, T s) { return {-s * a.y, s * a.x}; }

// Twiddle w_L^{sign e}; the table holds exp(-2 pi i e / L), sign -1 forward.
template <typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T> *tw, int e, int sign)
{
    cpx<T> t = tw[e];
    if (sign > 0) t.y = -t.y;
    return t;
}

template <typename T>
struct Consts {
    static constexpr T c3 = T(-0.5);
    static constexpr T s3 = T(0.86602540378443864676372317075294);
    static constexpr T c51 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    static constexpr T c52 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    static constexpr T s51 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    static constexpr T s52 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
};

// One Stockham pass over P lines of length L held in LDS (x -> y).
template <typename T>
__device__ void fft_pass(const cpx<T> *__restrict__ x, cpx<T> *__restrict__ y, int P, int L, int R,
                         int Ns, const cpx<T> *__restrict__ tw, int sign, int tid, int nthr)
{
    const int nb = L / R;           // butterflies per line
    const int tstep = L / (Ns * R); // table stride of w_{Ns R}
    const T sg = T(sign);
    if (R == 2 || R == 3 || R == 4 || R == 5) {
        for (int w = tid; w < P * nb; w += nthr) {
            const int c = w / nb, j = w - c * nb;
            const int k = j % Ns;
            const cpx<T> *xl = x + c * L;
            cpx<T> *yl = y + c * L;
            const int od = (j / Ns) * Ns * R + k;
            const int e1 = k * tstep;
            cpx<T> v0 = xl[j];
            if (R == 2) {
                cpx<T> v1 = cmul(xl[j + nb], twiddle(tw, e1, sign));
                yl[od] = cadd(v0, v1);
                yl[od + Ns] = csub(v0, v1);
            } else if (R == 3) {
                cpx<T> v1 = cmul(xl[j + nb], twiddle(tw, e1, sign));
                cpx<T> v2 = cmul(xl[j + 2 * nb], twiddle(tw, 2 * e1, sign));
                cpx<T> a = cadd(v1, v2), b = csub(v1, v2);
                cpx<T> t = {v0.x + Consts<T>::c3 * a.x, v0.y + Consts<T>::c3 * a.y};
                cpx<T> u = cmul_si(b, sg * Consts<T>::s3);
                yl[od] = cadd(v0, a);
                yl[od + Ns] = cadd(t, u);
                yl[od + 2 * Ns] = csub(t, u);
            } else if (R == 4) {
                cpx<T> v1 = cmul(xl[j + nb], twiddle(tw, e1, sign));
                cpx<T> v2 = cmul(xl[j + 2 * nb], twiddle(tw, 2 * e1, sign));
                cpx<T> v3 = cmul(xl[j + 3 * nb], twiddle(tw, 3 * e1, sign));
                cpx<T> s02 = cadd(v0, v2), d02 = csub(v0, v2);
                cpx<T> s13 = cadd(v1, v3), d13 = cmul_si(csub(v1, v3), sg);
                yl[od] = cadd(s02, s13);
                yl[od + Ns] = cadd(d02, d13);
                yl[od + 2 * Ns] = csub(s02, s13);
                yl[od + 3 * Ns] = csub(d02, d13);
            } else {  // R == 5
                cpx<T> v1 = cmul(xl[j + nb], twiddle(tw, e1, sign));
                cpx<T> v2 = cmul(xl[j + 2 * nb], twiddle(tw, 2 * e1, sign));
                cpx<T> v3 = cmul(xl[j + 3 * nb], twiddle(tw, 3 * e1, sign));
                cpx<T> v4 = cmul(xl[j + 4 * nb], twiddle(tw, 4 * e1, sign));
                cpx<T> a1 = cadd(v1, v4), b1 = csub(v1, v4);
                cpx<T> a2 = cadd(v2, v3), b2 = csub(v2, v3);
                const T c1 = Consts<T>::c51, c2 = Consts<T>::c52;
                const T s1 = Consts<T>::s51, s2 = Consts<T>::s52;
                cpx<T> t1 = {v0.x + c1 * a1.x + c2 * a2.x, v0.y + c1 * a1.y + c2 * a2.y};
                cpx<T> t2 = {v0.x + c2 * a1.x + c1 * a2.x, v0.y + c2 * a1.y + c1 * a2.y};
                cpx<T> u1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
                cpx<T> u2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
                u1 = cmul_si(u1, sg);
                u2 = cmul_si(u2, sg);
                yl[od] = {v0.x + a1.x + a2.x, v0.y + a1.y + a2.y};
                yl[od + Ns] = cadd(t1, u1);
                yl[od + 4 * Ns] = csub(t1, u1);
                yl[od + 2 * Ns] = cadd(t2, u2);
                yl[od + 3 * Ns] = csub(t2, u2);
            }
        }
    } else {
        // Direct DFT pass: output (j, q) = sum_r x[j + r nb] w_{Ns R}^{sign r (k + q Ns)}.
        for (int w = tid; w < P * L; w += nthr) {
            const int c = w / L, rem = w - c * L;
            const int q = rem / nb, j = rem - q * nb;
            const int k = j % Ns;
            const cpx<T> *xl = x + c * L;
            const int step = ((k + q * Ns) * tstep) % L;
            int e = 0;
            T ax = 0, ay = 0;
            for (int r = 0; r < R; ++r) {
                const cpx<T> t = twiddle(tw, e, sign);
                const cpx<T> v = xl[j + r * nb];
                ax += v.x * t.x - v.y * t.y;
                ay += v.x * t.y + v.y * t.x;
                e += step;
                if (e >= L) e -= L;
            }
            y[c * L + (j / Ns) * Ns * R + k + q * Ns] = {ax, ay};
        }
    }
}

// Full transform over P lines: returns the buffer holding the result.
template <typename T>
__device__ cpx<T> *fft_lines(cpx<T> *a, cpx<T> *b, int P, int L, int npass, const int *radix,
                             const cpx<T> *tw, int sign, int tid, int nthr)
{
    int Ns = 1;
    for (int p = 0; p < npass; ++p) {
        const int R = radix[p];
        fft_pass<T>(a, b, P, L, R, Ns, tw, sign, tid, nthr);
        __syncthreads();
        cpx<T> *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

}  // namespace scft
