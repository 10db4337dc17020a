// scft_kernels.h — device kernels of libscft (gfx950).
//
// Propagator: q_{s+1} = E . M (E . q_s) with E = exp(-DS w / 2) and
// M = F^-1 diag(K) F / (NX NY N) (SCFT.c:2474-2631).  K is separable,
// K = Kx Ky Kz, so M factors into one line operator per axis; each is one
// LDS-resident FFT -> K -> inverse FFT over a tile of grid lines.  The
// z operator with MASK 1 is the reference's Dirichlet transform: the line
// is odd-extended to N = 2(NZ+1) (SCFT.c:2551-2592), transformed, scaled by
// Kz on the extended grid, transformed back and its points 1..NZ kept.
#pragma once

#include <hip/hip_runtime.h>

#include "scft_fft.h"

namespace scft {

constexpr int kThreads = 256;
constexpr int kMaxTerms = 24;   // density terms (species blocks)
constexpr int kRedBlocks = 512; // partial sums per reduction

// One axis operator applied to every line of one propagator chain at one
// contour step.
template <typename T>
struct LineOp {
    const T *src;
    T *dst;
    const T *epre;    // Boltzmann factor applied on load (first axis), or null
    const T *epost;   // Boltzmann factor applied on store (last axis), or null
    const T *K;       // length L, K[m] of this axis
    const cpx<T> *tw; // length L, exp(-2 pi i e / L)
    int n, L, dirichlet;
    int nlines;       // real lines
    int c1, s0, s1;   // line base = (l / c1) * s1 + (l % c1) * s0
    int stride;       // element stride along the line
    int npass;
    int radix[kMaxPass];
};

template <typename T>
__device__ __forceinline__ int line_base(const LineOp<T> &op, int l)
{
    return (l / op.c1) * op.s1 + (l % op.c1) * op.s0;
}

// grid: (line groups, ops); block: kThreads; dynamic LDS: 2 * P * Lmax * sizeof(cpx<T>)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_lineop(const LineOp<T> *__restrict__ ops, int P)
{
    const LineOp<T> &op = ops[blockIdx.y];
    const int lines_per = 2 * P;
    const int g0 = blockIdx.x * lines_per;
    if (g0 >= op.nlines) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int L = op.L, n = op.n;
    cpx<T> *A = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *B = A + P * L;
    const int tid = threadIdx.x;
    const int nl = min(lines_per, op.nlines - g0);
    const int total = lines_per * n;
    const bool contiguous = (op.stride == 1);

    if (op.dirichlet) {
        for (int c = tid; c < P; c += kThreads) {
            A[c * L] = {T(0), T(0)};
            A[c * L + n + 1] = {T(0), T(0)};
        }
    }
    for (int e = tid; e < total; e += kThreads) {
        int r, t;
        if (contiguous) { r = e / n; t = e - r * n; }
        else { t = e / lines_per; r = e - t * lines_per; }
        T v = T(0);
        if (r < nl) {
            const int a = line_base(op, g0 + r) + t * op.stride;
            v = op.src[a];
            if (op.epre) v *= op.epre[a];
        }
        const int c = r >> 1;
        T *slot = reinterpret_cast<T *>(&A[c * L + (op.dirichlet ? t + 1 : t)]) + (r & 1);
        *slot = v;
        if (op.dirichlet) {
            T *mirror = reinterpret_cast<T *>(&A[c * L + L - 1 - t]) + (r & 1);
            *mirror = -v;
        }
    }
    __syncthreads();
    cpx<T> *X = fft_lines<T>(A, B, P, L, op.npass, op.radix, op.tw, -1, tid, kThreads);
    for (int e = tid; e < P * L; e += kThreads) {
        const T k = op.K[e % L];
        X[e].x *= k;
        X[e].y *= k;
    }
    __syncthreads();
    cpx<T> *Y = (X == A) ? B : A;
    X = fft_lines<T>(X, Y, P, L, op.npass, op.radix, op.tw, +1, tid, kThreads);
    const T scale = T(1) / T(L);
    for (int e = tid; e < total; e += kThreads) {
        int r, t;
        if (contiguous) { r = e / n; t = e - r * n; }
        else { t = e / lines_per; r = e - t * lines_per; }
        if (r >= nl) continue;
        const int c = r >> 1;
        const T *slot = reinterpret_cast<const T *>(&X[c * L + (op.dirichlet ? t + 1 : t)]) + (r & 1);
        const int a = line_base(op, g0 + r) + t * op.stride;
        T v = *slot * scale;
        if (op.epost) v *= op.epost[a];
        op.dst[a] = v;
    }
}

// ------------------------------------------------------------------ prep

struct PrepArgs {
    const double *w;     // 13*V
    int fields[13];      // fields whose Boltzmann factor is needed
    int nfields;
    double ds;
    int V;
    // dbk1 seed start (SCFT.c:2346-2367): mode 0 none, 1 seed box, 2 ones
    int seed_mode;
    int nx, ny, nz, delta;
};

template <typename T>
__global__ void __launch_bounds__(kThreads) k_prep(PrepArgs a, T *E, T *q0_dbk1)
{
    const int stride = gridDim.x * blockDim.x;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < a.V; idx += stride) {
        for (int i = 0; i < a.nfields; ++i) {
            const int f = a.fields[i];
            const double w = a.w[(size_t)f * a.V + idx];
            E[(size_t)f * a.V + idx] = T(exp(a.ds * (-w / 2.0)));
        }
        if (a.seed_mode == 2) {
            q0_dbk1[idx] = T(1);
        } else if (a.seed_mode == 1) {
            const int k = idx % a.nz;
            const int j = (idx / a.nz) % a.ny;
            const int i = idx / (a.nz * a.ny);
            const int D = a.delta;
            const bool in = i > (a.nx / 2 - D) && i < (a.nx / 2 + D) && j > (a.ny / 2 - D) &&
                            j < (a.ny / 2 + D) && k > (a.nz / 4 - D) && k < (a.nz / 4 + D);
            q0_dbk1[idx] = in ? T(1) : T(0);
        }
    }
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of up to 4 values; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double red[kThreads / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
    if (lane == 0)
        for (int i = 0; i < NV; ++i) red[wid][i] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) s += red[w][i];
            v[i] = s;
        }
}

// Q partial sums: grid (kRedBlocks, nspecies); ends[y] = q(NS_p - 1) slice.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_qsum(const T *const *ends, int V, double *partial)
{
    const T *q = ends[blockIdx.y];
    double v[1] = {0.0};
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads)
        v[0] += double(q[idx]);
    block_sum<1>(v);
    if (threadIdx.x == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// ------------------------------------------------ densities + F + RMSDev

template <typename T>
struct DenTerm {
    const T *q;
    const T *qd;
    int ns, start, stop;
    int field;   // 0..12
    int qslot;   // index into the Q array
    double frac; // volume factor (vol (1 - coating) or vol_s); <= 0: not applied
};

template <typename T>
struct DenArgs {
    DenTerm<T> term[kMaxTerms];
    int nterms;
    int V;
    int mask;
    int nphi;            // fields carrying density
    int phif[13];
    double chi[13][13];
    const double *Q;     // per qslot
    const double *w;     // 13*V
    const double *p;     // V
    double *tot;         // 13*V
    double *phitot;      // V
    double *partial;     // [2][gridDim.x]: H, (phi_tot-1)^2
};

template <typename T>
__device__ __forceinline__ double simpson(const DenTerm<T> &t, int V, int idx)
{
    const size_t sV = (size_t)V;
    const int top = t.ns - 1;
    auto qc = [&](int s) -> T { return t.q[(size_t)s * sV + idx] * t.qd[(size_t)(top - s) * sV + idx]; };
    if constexpr (sizeof(T) == 8) {
        // the reference's order: ends, then 4*odd, then 2*even (SCFT.c:2283-2297)
        double phi = double(qc(t.start)) + double(qc(t.stop));
        for (int s = t.start + 1; s <= t.stop - 1; s += 2) phi = phi + 4.0 * double(qc(s));
        for (int s = t.start + 2; s <= t.stop - 2; s += 2) phi = phi + 2.0 * double(qc(s));
        return phi;
    } else {
        // FP32 with Kahan-compensated summation
        float sum = qc(t.start) + qc(t.stop), comp = 0.0f;
        for (int s = t.start + 1; s <= t.stop - 1; ++s) {
            const float wgt = ((s - t.start) & 1) ? 4.0f : 2.0f;
            const float y = wgt * qc(s) - comp;
            const float tt = sum + y;
            comp = (tt - sum) - y;
            sum = tt;
        }
        return double(sum) - double(comp);
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_density(const DenArgs<T> *__restrict__ argp)
{
    const DenArgs<T> &a = *argp;
    const int V = a.V;
    double acc[2] = {0.0, 0.0};
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads) {
        double phi[13];
        for (int f = 0; f < 13; ++f) phi[f] = 0.0;
        bool touched[13];
        for (int f = 0; f < 13; ++f) touched[f] = false;
        for (int i = 0; i < a.nterms; ++i) {
            const DenTerm<T> &t = a.term[i];
            const double dseff = 1.0 / (t.ns - 1);
            double v = simpson<T>(t, V, idx) / 3.0;
            v = v * (dseff / a.Q[t.qslot]);
            if (t.frac > 0.0) v = t.frac * v;
            phi[t.field] = touched[t.field] ? phi[t.field] + v : v;
            touched[t.field] = true;
        }
        double tot = phi[0] + phi[1];
        for (int f = 2; f < 10; ++f) tot = tot + phi[f];
        if (a.mask)
            for (int f = 10; f < 13; ++f) tot = tot + phi[f];
        for (int f = 0; f < 13; ++f)
            if (touched[f]) a.tot[(size_t)f * V + idx] = phi[f];
        a.phitot[idx] = tot;
        // funFreeEnergy integrand (SCFT.c:1888-1910), pre-update w and p
        double h = 0.0;
        for (int i = 0; i < a.nphi; ++i)
            for (int j = i + 1; j < a.nphi; ++j) {
                const int f = a.phif[i], g = a.phif[j];
                h += a.chi[f][g] * phi[f] * phi[g];
            }
        double sphi = 1.0;
        for (int i = 0; i < a.nphi; ++i) {
            const int f = a.phif[i];
            h -= a.w[(size_t)f * V + idx] * phi[f];
            sphi -= phi[f];
        }
        h -= a.p[idx] * sphi;
        acc[0] += h;
        acc[1] += (tot - 1.0) * (tot - 1.0);
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        a.partial[blockIdx.x] = acc[0];
        a.partial[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ------------------------------------------------------------- update

struct UpdArgs {
    int V, nx, ny, nz;
    int nact;
    int act[13];
    double chi[13][13];
    double phibar[13];
    double g1, g2;
    double amp;          // dt * MF
    const double *noise; // draw order (k, j, i outer; active slot inner), or null
    double *w;           // 13*V
    double *p;           // V
    const double *tot;   // 13*V
    double *partial;     // [2][gridDim.x]: sum p_new, sum dw^2
};

// funUpdate per voxel (SCFT.c:2047-2113)
__global__ void __launch_bounds__(kThreads) k_update(const UpdArgs *__restrict__ argp)
{
    const UpdArgs &a = *argp;
    const int V = a.V;
    double acc[2] = {0.0, 0.0};
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads) {
        double w0[13], r[13], dW[13];
        const int n = a.nact;
        for (int i = 0; i < n; ++i) {
            const int f = a.act[i];
            w0[i] = a.w[(size_t)f * V + idx];
            r[i] = a.tot[(size_t)f * V + idx] - a.phibar[f];
        }
        const double pv = a.p[idx];
        for (int i = 0; i < n; ++i) {
            double d = -w0[i] + pv;
            for (int j = 0; j < n; ++j)
                if (j != i) d += a.chi[a.act[i]][a.act[j]] * r[j];
            dW[i] = d;
        }
        double corr = 0.0;
        if (n > 1)
            for (int i = 0; i < n; ++i)
                for (int j = i + 1; j < n; ++j) {
                    const double c = a.chi[a.act[i]][a.act[j]];
                    for (int h = 0; h < n; ++h)
                        if (h != i && h != j) corr += c * r[h];
                }
        size_t draw = 0;
        if (a.noise) {
            const int k = idx % a.nz, j = (idx / a.nz) % a.ny, i = idx / (a.nz * a.ny);
            draw = ((size_t)(k * a.ny + j) * a.nx + i) * n;
        }
        double psum = 0.0;
        for (int i = 0; i < n; ++i) {
            double wn = w0[i];
            for (int j = 0; j < n; ++j) wn += ((i == j) ? a.g1 : a.g2) * dW[j];
            if (a.noise) wn += a.amp * a.noise[draw + i];
            a.w[(size_t)a.act[i] * V + idx] = wn;
            psum += wn;
            acc[1] += (wn - w0[i]) * (wn - w0[i]);
        }
        const double pn = (psum + corr) / n;
        a.p[idx] = pn;
        acc[0] += pn;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        a.partial[blockIdx.x] = acc[0];
        a.partial[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ------------------------------------------------------------- finalize

struct Scalars {
    double Q[17];
    double H;        // F before the ln Q terms
    double F;
    double rms;
    double pmean;
    double dw_rms;
    double lnq_frac;
};

struct FinArgs {
    int V, nblocks, nq, nact;
    int qsp[17];          // species of each Q slot
    double lnq_coef[17];  // vol/alpha * c per Q slot (SCFT.c:1529-1555)
    const double *qpart;  // [nq][nblocks]
    const double *dpart;  // [2][nblocks]
    const double *upart;  // [2][nblocks]
    Scalars *out;
};

// Phase 1: Q from its partials (before the density kernel).
__global__ void k_qfinal(const FinArgs *__restrict__ argp)
{
    const FinArgs &a = *argp;
    const int s = blockIdx.x;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < a.nblocks; b += kThreads) v[0] += a.qpart[s * a.nblocks + b];
    block_sum<1>(v);
    if (threadIdx.x == 0) a.out->Q[a.qsp[s]] = v[0] / a.V;
}

// Phase 2: F, RMSDev, mean p, ||dw||.
__global__ void k_final(const FinArgs *__restrict__ argp)
{
    const FinArgs &a = *argp;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < a.nblocks; b += kThreads) {
        v[0] += a.dpart[b];
        v[1] += a.dpart[a.nblocks + b];
        v[2] += a.upart[b];
        v[3] += a.upart[a.nblocks + b];
    }
    block_sum<4>(v);
    if (threadIdx.x == 0) {
        Scalars *o = a.out;
        double frac = 0.0;
        for (int s = 0; s < a.nq; ++s) frac += a.lnq_coef[s] * log(o->Q[a.qsp[s]]);
        o->H = v[0] / a.V;
        o->lnq_frac = frac;
        o->F = o->H - frac;
        o->rms = sqrt(v[1] / a.V);
        o->pmean = v[2] / a.V;
        o->dw_rms = a.nact > 0 ? sqrt(v[3] / ((double)a.nact * a.V)) : 0.0;
    }
}

__global__ void __launch_bounds__(kThreads) k_pshift(double *p, int V, const Scalars *s)
{
    const double m = s->pmean;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads) p[idx] -= m;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_convert(const double *src, T *dst, int n)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) dst[i] = T(src[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_widen(const T *src, double *dst, int n)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) dst[i] = double(src[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill(T *dst, T v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) dst[i] = v;
}

}  // namespace scft
