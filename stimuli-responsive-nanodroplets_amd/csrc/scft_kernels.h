// scft_kernels.h — device kernels of libscft (gfx950).
//
// Propagator: q_{s+1} = E . M (E . q_s) with E = exp(-DS w / 2) and
// M = F^-1 diag(K) F / (NX NY N) (SCFT.c:2474-2631).  K is separable,
// K = Kx Ky Kz, so M factors into one real line operator per axis, applied
// to every grid line of that axis, z first.  Slices are stored without the
// factor of the step that produced them; the first axis of the next step
// applies it together with its own (see pre_factor), and the readers of
// slices (Q, contour sums, slice read-back) apply it on read.
//   - k_zdense_mfma2 / k_zdense_mfma / k_zdense: the z operator as an
//     NZ x NZ matrix.  M commutes with the reversal k -> NZ-1-k, so it
//     splits into a symmetric and an antisymmetric half of (NZ/2)^2 each:
//     NZ^2/2 FMA per line on FP64 MFMA (VALU for NZ > 128), against the
//     FFT of length 2(NZ+1) = 2*3*43 (NZ = 128) whose prime factor 43 costs
//     ~6x that.  The autotuner picks per size.
//   - k_fftreg<P>: periodic lines of length 16 P (128, 256) as a four-step
//     FFT in registers; k_fftreg2<B>: the same for 16 B (192, 320), 16
//     lanes of B points; k_fftmix: the lengths 192, 256 and 320 of a batch
//     of uneven NX in one launch; k_fftwave: other compile-time lengths, one
//     wavefront per line; k_lineop: any length, LDS Stockham passes, and the
//     reference's Dirichlet transform (odd extension to N = 2(NZ+1),
//     SCFT.c:2551-2592) when the dense z operator is not used.
// All transform arithmetic is FP64; propagator slices (and the scratch
// between axes) are stored as T (double, or float for the FP32-split
// mode).  Every launch covers every live chain of every active instance of
// the batch.  Around the propagator: k_prep (Boltzmann factors, seed
// start), k_qsum / k_density (Q and the contour sums), k_update / k_final
// (the field update, F, RMSDev) and k_rng_* (the reference's noise stream
// drawn on the device).
#pragma once

#include <hip/hip_runtime.h>

#include "scft_fft.h"

namespace scft {

constexpr int kThreads = 256;
constexpr int kMaxTerms = 24;   // density terms (species blocks)
constexpr int kRedBlocks = 1024; // partial sums (workgroups) per reduction and instance
constexpr int kDenseTM = 64;    // rows per k_zdense workgroup
constexpr int kDenseHMax = 64;  // k_zdense handles NZ <= 128

// Global-memory accessors.  Pointers reach the kernels inside argument
// blocks, so the compiler cannot tell their address space and would emit
// flat accesses, which also count against the LDS/scalar wait counter and
// serialise every load with the LDS traffic around it.  Offsets are 32-bit
// element offsets from a per-slice (or per-table) base: the base stays in
// SGPRs and the byte offset in one VGPR (saddr + voffset addressing), with
// no 64-bit address arithmetic per access.  Slices are limited to 2^28
// points, so offset * sizeof(T) < 2^32.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gaddr(const T *p, unsigned i)
{
    return (const __attribute__((address_space(1))) T *)((const char *)p + i * (unsigned)sizeof(T));
}

template <typename T>
__device__ __forceinline__ T ldg(const T *p, unsigned i)
{
    return *gaddr(p, i);
}

// p[i] as a streaming (non-temporal) load: data read once per launch.
template <typename T>
__device__ __forceinline__ T ldg_nt(const T *p, unsigned i)
{
    return __builtin_nontemporal_load(gaddr(p, i));
}

template <typename T>
__device__ __forceinline__ void stg(T *p, unsigned i, T v)
{
    *(__attribute__((address_space(1))) T *)((char *)p + i * (unsigned)sizeof(T)) = v;
}

template <typename T>
struct Pair {
    T x, y;
};

// p[i], p[i + 1] in one load (i even, p aligned to two elements).
template <typename T>
__device__ __forceinline__ Pair<T> ldg_pair(const T *p, unsigned i)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = *(const __attribute__((address_space(1))) v2 *)((const char *)p + i * (unsigned)sizeof(T));
    return Pair<T>{v.x, v.y};
}

// The same as a streaming (non-temporal) load: for data read once per
// launch, so that it does not displace reused operands from the caches.
template <typename T>
__device__ __forceinline__ Pair<T> ldg_pair_nt(const T *p, unsigned i)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load((const __attribute__((address_space(1))) v2 *)((const char *)p +
                                                                                         i * (unsigned)sizeof(T)));
    return Pair<T>{v.x, v.y};
}

// p[i], p[i + 1] in one store (i even, p aligned to two elements).
template <typename T>
__device__ __forceinline__ void stg_pair(T *p, unsigned i, T a, T b)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    *(__attribute__((address_space(1))) v2 *)((char *)p + i * (unsigned)sizeof(T)) = v2{a, b};
}

// A propagator slice value stored in the slice type T.  FP64 slices keep
// the value; FP32 slices round it without bias when the op carries a
// rounding key (rkey != 0): a hashed 29-bit integer (from the op's key and
// the element offset) is added to the FP64 bits and the 29 bits below an
// FP32 mantissa are cleared, so the value goes to one of its two FP32
// neighbours with probability set by the distance.  Round-to-nearest maps
// mirror-image values that differ by less than half an FP32 step to the
// same value at every step, so an asymmetric perturbation below that size
// is erased rather than carried; FP64 keeps it, and a symmetric state that
// is unstable in FP64 then looks stable in FP32.  Unbiased rounding keeps
// each value right on average and leaves such a perturbation a chance to
// grow.  Deterministic: the same chain, step and element always draw the
// same bits.  The last axis of a step (the stored slice) rounds this way;
// the scratch between axes rounds to nearest.
template <typename T>
__device__ __forceinline__ T store_round(double v, unsigned rkey, unsigned i);

template <>
__device__ __forceinline__ double store_round<double>(double v, unsigned, unsigned)
{
    return v;
}

template <>
__device__ __forceinline__ float store_round<float>(double v, unsigned rkey, unsigned i)
{
    if (!rkey) return float(v);
    unsigned h = (i ^ rkey) * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    const unsigned long long b =
        ((unsigned long long)__double_as_longlong(v) + (unsigned long long)(h >> 3)) & ~0x1FFFFFFFull;
    return float(__longlong_as_double((long long)b));   // exact for normal FP32 values
}

// ------------------------------------------------------------ FFT line op

struct LineOp {
    const void *src;
    void *dst;
    const double *epre;  // Boltzmann factor applied on load (first axis), or null
    const double *epre2; // second factor on load (see pre_factor), or null
    const double *epost; // Boltzmann factor applied on store, or null
    const double *K;     // length L
    const cpx<double> *tw;
    int n, L, dirichlet;
    int nlines;          // real lines
    int c1, s0, s1;      // line base = (l / c1) * s1 + (l % c1) * s0
    int stride;          // element stride along the line
    int pairs;           // lines 2i, 2i+1 are adjacent and 2-element aligned (nlines even)
    int npass;
    int radix[kMaxPass];
    unsigned rkey;       // unbiased-rounding key of FP32 stores (store_round), 0: round to nearest
};

__device__ __forceinline__ int line_base(const LineOp &op, int l)
{
    return (l / op.c1) * op.s1 + (l % op.c1) * op.s0;
}

// Factor applied on load by the first axis of a contour step.  Slices are
// stored without their post-step Boltzmann factor, so the first axis of the
// next step applies E_f(t) E_f(t-1): one load squared inside a block
// (e2 == e1), two loads at a block boundary, E alone on the start slice.
__device__ __forceinline__ double pre_factor(const double *e1, const double *e2, unsigned a)
{
    const double x = ldg(e1, a);
    if (!e2) return x;
    return (e2 == e1) ? x * x : x * ldg(e2, a);
}

// grid: (line groups, ops); dynamic LDS: 2 * P * Lmax * 16 bytes
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kThreads) k_lineop(const LineOp *__restrict__ ops, int P)
{
    using T = double;
    const LineOp &op = ops[blockIdx.y];
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
    const Tin *__restrict__ src = static_cast<const Tin *>(op.src);
    Tout *__restrict__ dst = static_cast<Tout *>(op.dst);

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
            v = T(src[a]);
            if (op.epre) v *= pre_factor(op.epre, op.epre2, a);
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
        dst[a] = store_round<Tout>(v, op.rkey, unsigned(a));
    }
}

// --------------------------------------------------------- dense z op

struct DenseOp {
    const void *src;
    void *dst;
    const double *epre, *epre2, *epost;
    const float *e2f;    // E^2 of epre's field in FP32 (inside a block, FP32-split), or null
    const double *MsT;   // [hs][ldm]: MsT[j][k] = Ms[k][j]
    const double *MaT;   // [ha][ldm]
    int n, hs, ha, ldm;
    int nrows;
};

// grid: (ceil(rows / kDenseTM), ops); 256 threads = 16 row groups x 16 k groups,
// each thread 4 rows x 4 k of both halves.
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kThreads) k_zdense(const DenseOp *__restrict__ ops)
{
    const DenseOp &op = ops[blockIdx.y];
    const int row0 = blockIdx.x * kDenseTM;
    if (row0 >= op.nrows) return;
    constexpr int LDS = kDenseHMax + 1;   // odd stride: the 4 row groups of a wave hit different banks
    __shared__ double S[kDenseTM * LDS];
    __shared__ double Aa[kDenseTM * LDS];
    const int n = op.n, hs = op.hs, ha = op.ha;
    const Tin *__restrict__ src = static_cast<const Tin *>(op.src);
    Tout *__restrict__ dst = static_cast<Tout *>(op.dst);
    const int tid = threadIdx.x;
    // symmetric / antisymmetric parts of each row (E applied on load)
    for (int e = tid; e < kDenseTM * hs; e += kThreads) {
        const int r = e / hs, j = e - r * hs;
        const int row = row0 + r;
        double s = 0.0, a = 0.0;
        if (row < op.nrows) {
            const size_t base = (size_t)row * n;
            double x1 = double(src[base + j]);
            if (op.epre) x1 *= pre_factor(op.epre, op.epre2, base + j);
            if (j < ha) {
                double x2 = double(src[base + n - 1 - j]);
                if (op.epre) x2 *= pre_factor(op.epre, op.epre2, base + n - 1 - j);
                s = x1 + x2;
                a = x1 - x2;
            } else {
                s = x1;   // middle point of an odd line
            }
        }
        S[r * LDS + j] = s;
        if (j < ha) Aa[r * LDS + j] = a;
    }
    __syncthreads();
    const int rg = tid >> 4, kg = tid & 15;
    for (int k0 = kg * 4; k0 < hs; k0 += 64) {
        double u[4][4], v[4][4];
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 4; ++c) { u[i][c] = 0.0; v[i][c] = 0.0; }
        for (int j = 0; j < hs; ++j) {
            double s[4];
            for (int i = 0; i < 4; ++i) s[i] = S[(rg * 4 + i) * LDS + j];
            const double *m = op.MsT + (size_t)j * op.ldm + k0;
            const double m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
            for (int i = 0; i < 4; ++i) {
                u[i][0] += s[i] * m0;
                u[i][1] += s[i] * m1;
                u[i][2] += s[i] * m2;
                u[i][3] += s[i] * m3;
            }
        }
        for (int j = 0; j < ha; ++j) {
            double a[4];
            for (int i = 0; i < 4; ++i) a[i] = Aa[(rg * 4 + i) * LDS + j];
            const double *m = op.MaT + (size_t)j * op.ldm + k0;
            const double m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
            for (int i = 0; i < 4; ++i) {
                v[i][0] += a[i] * m0;
                v[i][1] += a[i] * m1;
                v[i][2] += a[i] * m2;
                v[i][3] += a[i] * m3;
            }
        }
        for (int i = 0; i < 4; ++i) {
            const int row = row0 + rg * 4 + i;
            if (row >= op.nrows) continue;
            const size_t base = (size_t)row * n;
            for (int c = 0; c < 4; ++c) {
                const int k = k0 + c;
                if (k >= hs) continue;
                double y1 = u[i][c] + v[i][c];
                if (op.epost) y1 *= op.epost[base + k];
                dst[base + k] = Tout(y1);
                if (k < ha) {
                    double y2 = u[i][c] - v[i][c];
                    if (op.epost) y2 *= op.epost[base + n - 1 - k];
                    dst[base + n - 1 - k] = Tout(y2);
                }
            }
        }
    }
}

// ------------------------------------------------ dense z op on MFMA (FP64)
//
// v_mfma_f64_16x16x4f64 lane layout on gfx950 (checked on the device):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register r holds D[l / 16 + 4 r][l % 16]
// U = S Ms^T and V = A Ma^T for 16 rows per wave: the symmetric/antisymmetric
// row parts are formed straight into A-fragment registers (E applied on
// load), Ms^T / Ma^T are staged once per workgroup in LDS, HSP = hs padded
// to a multiple of 16 (zero rows/columns).

typedef double d4_t __attribute__((ext_vector_type(4)));

// The wave owns its line: LDS operations of one wavefront are performed in
// order, so a wavefront-scope fence (a compiler barrier) orders the passes.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS row stride of the staged Ms^T / Ma^T: a half-wave reads fragment rows
// t and NT + t; their offset NT*LDM doubles must be 16 (mod 32) doubles so
// the two 16-double runs fall into disjoint bank halves.
constexpr int mfma_ldm(int hsp)
{
    const int nt = hsp / 4;
    for (int l = hsp; l < hsp + 64; ++l)
        if ((nt * l) % 32 == 16) return l;
    return hsp + 1;
}

// Wave tiles of 16 rows.  A tile is loaded with coalesced row loads (lanes
// along z), staged in LDS with the Boltzmann factors applied, and the
// symmetric/antisymmetric A fragments are read from LDS: the MFMA operand
// layout pairs lanes with rows, so fragments loaded straight from global
// memory would gather 16 rows per instruction and re-fetch lines.  The
// workgroup stages Ms^T / Ma^T once and its waves then stride over the
// launch's tiles independently; the next tile's loads are issued before the
// current tile's products, so one workgroup per CU keeps HBM busy.
constexpr int kMfmaWaves = 4;

template <int HSP>
struct MfmaTile {
    static constexpr int NC = (2 * HSP + 63) / 64;   // 64-lane chunks per row
    static constexpr int NE = 16 * NC;               // values per lane per tile
};

// Loads of one wave tile held in registers until the tile is staged.  The
// z operator is the first axis of a step, so epre is always set.  The
// second factor of a block-boundary step (one step per block) is read at
// staging time, to keep the prefetch within the registers.
template <typename Tin, int HSP>
struct TileRegs {
    Tin x[MfmaTile<HSP>::NE];
    double e[MfmaTile<HSP>::NE];
};

// Element u of a lane's tile share: row u / NC, column lane + 64 (u % NC).
// Loads are unconditional at clamped addresses (row 0 / column 0 of the op
// for positions past its end) so that they issue back to back; staging
// zeroes those positions.
template <int NC>
__device__ __forceinline__ unsigned mfma_addr(int u, int lane, int rbase, int nr, int n)
{
    const int r = u / NC, c = lane + 64 * (u % NC);
    return unsigned(rbase + (r < nr ? r : 0)) * unsigned(n) + unsigned(c < n ? c : 0);
}

// Tiles past the launch's end reload the last tile: the pipeline stays
// free of branches, which would make the compiler wait for the loads early.
template <typename Tin, int HSP>
__device__ __forceinline__ void mfma_fetch(const DenseOp *__restrict__ ops, int tpo, int t, int ntiles, int lane,
                                           TileRegs<Tin, HSP> &R)
{
    constexpr int NC = MfmaTile<HSP>::NC, NE = MfmaTile<HSP>::NE;
    t = min(t, ntiles - 1);
    const DenseOp &op = ops[t / tpo];
    const int rbase = (t % tpo) * 16, n = op.n;
    const int nr = min(16, op.nrows - rbase);
    const Tin *src = static_cast<const Tin *>(op.src);
#pragma unroll
    for (int u = 0; u < NE; ++u) R.x[u] = ldg(src, mfma_addr<NC>(u, lane, rbase, nr, n));
#pragma unroll
    for (int u = 0; u < NE; ++u) R.e[u] = ldg(op.epre, mfma_addr<NC>(u, lane, rbase, nr, n));
}

// Stage tile t from the registers into the wave's LDS rows with the
// Boltzmann factors applied (rows past the op's end are zero).
template <typename Tin, int HSP, int XS>
__device__ __forceinline__ void mfma_stage(const DenseOp *__restrict__ ops, int tpo, int t, int ntiles, int lane,
                                           const TileRegs<Tin, HSP> &R, double *Xw)
{
    constexpr int NC = MfmaTile<HSP>::NC, NE = MfmaTile<HSP>::NE;
    t = min(t, ntiles - 1);
    const DenseOp &op = ops[t / tpo];
    const int rbase = (t % tpo) * 16, n = op.n;
    const int nr = min(16, op.nrows - rbase);
    const int mode = !op.epre2 ? 1 : (op.epre2 == op.epre ? 2 : 3);   // E, E^2, E E2
    if (mode == 3) {
        double f[NE];
#pragma unroll
        for (int u = 0; u < NE; ++u) f[u] = ldg(op.epre2, mfma_addr<NC>(u, lane, rbase, nr, n));
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int r = u / NC, c = lane + 64 * (u % NC);
            if (c < n) Xw[r * XS + c] = r < nr ? double(R.x[u]) * (R.e[u] * f[u]) : 0.0;
        }
    } else {
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int r = u / NC, c = lane + 64 * (u % NC);
            const double fac = mode == 2 ? R.e[u] * R.e[u] : R.e[u];
            if (c < n) Xw[r * XS + c] = r < nr ? double(R.x[u]) * fac : 0.0;
        }
    }
}

// Per wave and tile: build the A fragments of tile t from LDS, stage tile
// t+1 (loaded during the previous tile's products) into the freed rows,
// issue the loads of tile t+2, run the products, store D straight from the
// accumulators.  The loads get a whole tile of products to land.
template <typename Tin, typename Tout, int HSP>
__global__ void __launch_bounds__(64 * kMfmaWaves) k_zdense_mfma(const DenseOp *__restrict__ ops, int ntiles, int tpo)
{
    constexpr int NT = HSP / 4;    // K steps
    constexpr int NK = HSP / 16;   // output column tiles
    constexpr int LDM = mfma_ldm(HSP);
    constexpr int XS = 2 * HSP + 1;   // staged row stride (doubles), odd: rows hit different banks
    __shared__ double Ms[HSP * LDM];
    __shared__ double Ma[HSP * LDM];
    __shared__ double X[kMfmaWaves * 16 * XS];
    // every op of a launch shares one operator (same NZ, LZ and NS)
    const int n = ops[0].n, hs = ops[0].hs, ha = ops[0].ha;
    const int tid = threadIdx.x;
    for (int e = tid; e < HSP * HSP; e += 64 * kMfmaWaves) {
        const int j = e / HSP, k = e - j * HSP;
        Ms[j * LDM + k] = (j < hs && k < hs) ? ldg(ops[0].MsT, unsigned(j * ops[0].ldm + k)) : 0.0;
        Ma[j * LDM + k] = (j < ha && k < ha) ? ldg(ops[0].MaT, unsigned(j * ops[0].ldm + k)) : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform: tile indices stay scalar
    const int li = lane & 15, lj = lane >> 4;
    double *Xw = X + wave * 16 * XS;
    const int wstride = gridDim.x * kMfmaWaves;
    int t = blockIdx.x * kMfmaWaves + wave;
    if (t >= ntiles) return;   // no workgroup barrier follows
    TileRegs<Tin, HSP> R;
    mfma_fetch<Tin, HSP>(ops, tpo, t, ntiles, lane, R);
    mfma_stage<Tin, HSP, XS>(ops, tpo, t, ntiles, lane, R, Xw);
    mfma_fetch<Tin, HSP>(ops, tpo, t + wstride, ntiles, lane, R);
    for (; t < ntiles; t += wstride) {
        wave_sync();   // tile t staged
        double sf[NT], af[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int j = 4 * k + lj;
            double sv = 0.0, av = 0.0;
            if (j < ha) {
                const double x1 = Xw[li * XS + j], x2 = Xw[li * XS + n - 1 - j];
                sv = x1 + x2;
                av = x1 - x2;
            } else if (j < hs) {
                sv = Xw[li * XS + j];   // middle point of an odd line
            }
            sf[k] = sv;
            af[k] = av;
        }
        wave_sync();   // every lane has read its fragments
        mfma_stage<Tin, HSP, XS>(ops, tpo, t + wstride, ntiles, lane, R, Xw);
        mfma_fetch<Tin, HSP>(ops, tpo, t + 2 * wstride, ntiles, lane, R);
        d4_t U[NK], W[NK];
#pragma unroll
        for (int kt = 0; kt < NK; ++kt) {
            U[kt] = d4_t{0.0, 0.0, 0.0, 0.0};
            W[kt] = d4_t{0.0, 0.0, 0.0, 0.0};
        }
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const double *ms = Ms + (4 * k + lj) * LDM + li;
            const double *ma = Ma + (4 * k + lj) * LDM + li;
#pragma unroll
            for (int kt = 0; kt < NK; ++kt) {
                U[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sf[k], ms[16 * kt], U[kt], 0, 0, 0);
                W[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[k], ma[16 * kt], W[kt], 0, 0, 0);
            }
        }
        // D layout: register r of lane l is row l/16 + 4r, column l%16 of the
        // tile: 16 lanes store 16 consecutive z points of one row
        const DenseOp &op = ops[t / tpo];
        const int rbase = (t % tpo) * 16;
        const int nr = min(16, op.nrows - rbase);
        Tout *dst = static_cast<Tout *>(op.dst);
#pragma unroll
        for (int kt = 0; kt < NK; ++kt) {
            const int k = 16 * kt + li;
            if (k >= hs) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = lj + 4 * r;
                if (rr >= nr) continue;
                const unsigned row = unsigned(rbase + rr) * unsigned(n);
                const double u = U[kt][r], v = W[kt][r];
                double y = u + v;
                if (op.epost) y *= ldg(op.epost, row + k);
                stg(dst, row + k, Tout(y));
                if (k < ha) {
                    double y2 = u - v;
                    if (op.epost) y2 *= ldg(op.epost, row + n - 1 - k);
                    stg(dst, row + n - 1 - k, Tout(y2));
                }
            }
        }
    }
}

// Even NZ (hs == ha): two wavefronts per SIMD.  The staging rows are
// halved (8 at a time) so that eight waves and Ms^T / Ma^T fit in LDS; each
// lane loads adjacent point pairs, rows past the op's end load a valid row
// and are never stored, so the staging needs no masks.  The K index is
// permuted (j = lj NT + k): a lane's fragment values are consecutive z
// points and the two fragment rows a half-wave reads from Ms^T / Ma^T fall
// into disjoint bank halves (mfma_ldm).
constexpr int kMfma2Waves = 8;

template <int HSP>
struct Mfma2 {
    static constexpr int NT = HSP / 4, NK = HSP / 16, LDM = mfma_ldm(HSP);
    static constexpr int XS = 2 * HSP + 2;   // even: rows stay 16-byte aligned
};

// Ms^T / Ma^T of the op into LDS, zero-padded to HSP x HSP (row stride LDM).
template <int HSP>
__device__ __forceinline__ void mfma2_stage_ops(const DenseOp &op, double *Ms, double *Ma, int tid, int nthreads)
{
    constexpr int LDM = Mfma2<HSP>::LDM;
    const int hs = op.hs;
    for (int e = tid; e < HSP * HSP; e += nthreads) {
        const int j = e / HSP, k = e - j * HSP;
        const bool in = j < hs && k < hs;
        Ms[j * LDM + k] = in ? ldg(op.MsT, unsigned(j * op.ldm + k)) : 0.0;
        Ma[j * LDM + k] = in ? ldg(op.MaT, unsigned(j * op.ldm + k)) : 0.0;
    }
}

// The loads of one 16-row tile: two points per lane from each row of the
// slice and of the step's factor, held until the tile is staged.  An FP32
// E^2 pair (mode 4) is held as its raw 8 bytes in ea[r].x and widened at
// staging, so that every mode issues the same loads in the same order.
template <typename Tin>
struct Z2Regs {
    Pair<Tin> xa[16];
    Pair<double> ea[16];
};

// factor: E alone (start slice), E^2 inside a block (FP64 E squared, or
// the FP32 E^2 array: mode 4), E E2 at a block boundary
__device__ __forceinline__ int mfma2_mode(const DenseOp &op)
{
    return !op.epre2 ? 1 : (op.epre2 == op.epre ? (op.e2f ? 4 : 2) : 3);
}

// E2F = false: the caller has no FP32 E^2 factor (one load path).
template <typename Tin, int HSP, bool E2F = true>
__device__ __forceinline__ void mfma2_load(const DenseOp &op, int rbase, int lane, Z2Regs<Tin> &R)
{
    const int n = op.n;
    const int c = 2 * lane;
    const int cc = c + 1 < n ? c : 0;       // load column (pairs past the row read column 0)
    const int nr = min(16, op.nrows - rbase);
    const Tin *src = static_cast<const Tin *>(op.src);
    const unsigned o0 = unsigned(rbase) * unsigned(n) + unsigned(cc);
    if (E2F && mfma2_mode(op) == 4) {
        const double *e2 = reinterpret_cast<const double *>(op.e2f);   // a float pair per 8 bytes
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const unsigned a = o0 + unsigned(r < nr ? r : 0) * unsigned(n);
            R.xa[r] = ldg_pair_nt(src, a);   // the slice is read once; the factors are reused every step
            R.ea[r] = Pair<double>{ldg(e2, a >> 1), 0.0};   // a is even
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const unsigned a = o0 + unsigned(r < nr ? r : 0) * unsigned(n);
            R.xa[r] = ldg_pair_nt(src, a);
            R.ea[r] = ldg_pair(op.epre, a);
        }
    }
}

// Stage a loaded tile into the wave's 8-row LDS rows with its factor
// applied, in two halves, and read each lane's symmetric / antisymmetric
// A fragments (sf, af) back.
template <typename Tin, int HSP>
__device__ __forceinline__ void mfma2_stage(const DenseOp &op, int rbase, const Z2Regs<Tin> &R, double *Xw, int lane,
                                            double (&sf)[HSP / 4], double (&af)[HSP / 4])
{
    constexpr int NT = Mfma2<HSP>::NT, XS = Mfma2<HSP>::XS;
    const int n = op.n;
    const int li = lane & 15, lj = lane >> 4;
    const int c = 2 * lane;                 // this lane's point pair in a row
    const bool cl = c < 2 * HSP;            // lanes that own staged columns
    const int cc = c + 1 < n ? c : 0;
    const int nr = min(16, op.nrows - rbase);
    const int mode = mfma2_mode(op);
    const unsigned o0 = unsigned(rbase) * unsigned(n) + unsigned(cc);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        wave_sync();   // the previous reads of the staging rows are done
        if (cl) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int Rr = 8 * h + r;
                double e0 = R.ea[Rr].x, e1 = R.ea[Rr].y;
                if (mode == 4) {
                    const long long bits = __double_as_longlong(R.ea[Rr].x);
                    e0 = double(__int_as_float(int(bits & 0xFFFFFFFFll)));
                    e1 = double(__int_as_float(int(bits >> 32)));
                } else if (mode == 2) {
                    e0 *= e0;
                    e1 *= e1;
                } else if (mode == 3) {   // one step per block: load the second factor here
                    const Pair<double> f = ldg_pair(op.epre2, o0 + unsigned(Rr < nr ? Rr : 0) * unsigned(n));
                    e0 *= f.x;
                    e1 *= f.y;
                }
                double2 v;
                v.x = double(R.xa[Rr].x) * e0;
                v.y = double(R.xa[Rr].y) * e1;
                *reinterpret_cast<double2 *>(&Xw[r * XS + c]) = v;
            }
        }
        wave_sync();
        if ((li >> 3) == h) {
            const double *row = Xw + (li & 7) * XS;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                const int j = lj * NT + k;   // j >= hs meets zero rows of Ms^T / Ma^T
                const double x1 = row[j], x2 = row[max(n - 1 - j, 0)];
                sf[k] = x1 + x2;
                af[k] = x1 - x2;
            }
        }
    }
}

// The products of a staged tile and its stores to dst (the op's output,
// or a caller's scratch with the same layout).  Ms/Ma are the staged
// operator halves, Xw the wave's staging rows (reused for the FP32 rows).
// NROW > 0: the caller guarantees n = NROW (a multiple of 16, at most
// 2 HSP) and whole 16-row tiles with FP32 output, so the row-store path is
// the only one compiled, with a fixed number of stores.
template <typename Tout, int HSP, int NROW = 0>
__device__ __forceinline__ void mfma2_mma_store(const DenseOp &op, int rbase, Tout *dst, const double *Ms,
                                                const double *Ma, double *Xw, int lane, const double (&sf)[HSP / 4],
                                                const double (&af)[HSP / 4])
{
    constexpr int NT = Mfma2<HSP>::NT, NK = Mfma2<HSP>::NK, LDM = Mfma2<HSP>::LDM;
    const int n = op.n, hs = op.hs;   // n = 2 hs
    const int li = lane & 15, lj = lane >> 4;
    const int nr = min(16, op.nrows - rbase);
    d4_t U[NK], W[NK];
#pragma unroll
    for (int kt = 0; kt < NK; ++kt) {
        U[kt] = d4_t{0.0, 0.0, 0.0, 0.0};
        W[kt] = d4_t{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const double *ms = Ms + (lj * NT + k) * LDM + li;
        const double *ma = Ma + (lj * NT + k) * LDM + li;
#pragma unroll
        for (int kt = 0; kt < NK; ++kt) {
            U[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sf[k], ms[16 * kt], U[kt], 0, 0, 0);
            W[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[k], ma[16 * kt], W[kt], 0, 0, 0);
        }
    }
    // D layout: register r of lane l is row l/16 + 4r, output point l%16
    // of the column tile; the mirrored half goes to n-1-k
    static_assert(NROW == 0 || (sizeof(Tout) == 4 && NROW % 16 == 0 && NROW <= 2 * HSP), "row stores: FP32 rows");
    if constexpr (sizeof(Tout) == 4 && (NROW > 0 || HSP == 64 || HSP == 32)) {
        if (NROW > 0 || (n == 2 * HSP && nr == 16)) {
            // FP32 rows of NR points: the tile's results go through the wave's
            // staging rows (16 rows at a pitch of 2 HSP floats, the 16-byte
            // chunks of row rr permuted by 4 (rr & 3), so that the four rows
            // one write phase touches fall into different banks) and out as
            // whole rows, 16 bytes per lane, instead of 4-byte stores of
            // 64-byte runs
            constexpr int NR = NROW > 0 ? NROW : 2 * HSP;
            constexpr int RC = NR / 4;   // 16-byte chunks per row
            float *Yw = reinterpret_cast<float *>(Xw);
            wave_sync();   // every lane has read its fragments from the staging rows
#pragma unroll
            for (int kt = 0; kt < NK; ++kt) {
                const int k = 16 * kt + li;
                if (NR < 2 * HSP && k >= NR / 2) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rr = lj + 4 * r, f = 4 * (rr & 3);
                    const double u = U[kt][r], v = W[kt][r];
                    const int k2 = NR - 1 - k;
                    Yw[rr * 2 * HSP + ((((k >> 2) ^ f)) << 2) + (k & 3)] = Tout(u + v);
                    Yw[rr * 2 * HSP + ((((k2 >> 2) ^ f)) << 2) + (k2 & 3)] = Tout(u - v);
                }
            }
            wave_sync();
            typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int it = 0; it < NR / 16; ++it) {
                const int idx = 64 * it + lane, rr = idx / RC, c = idx % RC;
                const f4 x = *reinterpret_cast<const f4 *>(Yw + rr * 2 * HSP + ((c ^ (4 * (rr & 3))) << 2));
                *(__attribute__((address_space(1))) f4 *)((char *)dst +
                                                          (unsigned(rbase + rr) * unsigned(NR) + unsigned(4 * c)) * 4u) = x;
            }
            return;
        }
    }
#pragma unroll
    for (int kt = 0; kt < NK; ++kt) {
        const int k = 16 * kt + li;
        if (k >= hs) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rr = lj + 4 * r;
            if (rr >= nr) continue;
            const unsigned row = unsigned(rbase + rr) * unsigned(n);
            const double u = U[kt][r], v = W[kt][r];
            stg(dst, row + k, Tout(u + v));
            stg(dst, row + n - 1 - k, Tout(u - v));
        }
    }
}

// One 16-row tile (rows rbase..rbase+15 of op) of the z operator, result to
// dst: loads, staging, products and stores in sequence.
template <typename Tin, typename Tout, int HSP>
__device__ __forceinline__ void mfma2_tile(const DenseOp &op, int rbase, Tout *dst, const double *Ms,
                                           const double *Ma, double *Xw, int lane)
{
    Z2Regs<Tin> R;
    mfma2_load<Tin, HSP>(op, rbase, lane, R);
    double sf[HSP / 4], af[HSP / 4];
    mfma2_stage<Tin, HSP>(op, rbase, R, Xw, lane, sf, af);
    mfma2_mma_store<Tout, HSP>(op, rbase, dst, Ms, Ma, Xw, lane, sf, af);
}

// The 16-row tiles of a z launch.  A batch of uneven NX has ops of
// different row counts: the ops come grouped by row count, group g's from
// op op0[g] on with tpo[g] tiles each and its tiles numbered from t0[g], so
// that no tile lies past its op's last row (with one tile count for every
// op, the shorter ops' spare tiles ran the whole tile for nothing).
constexpr int kZGroups = 4;

struct ZTiles {
    int ng;
    int op0[kZGroups], tpo[kZGroups], t0[kZGroups];
};

// Tile t: op index and first row (wave-uniform).  GR false: one group (every
// op with the same row count), the plain division.
template <bool GR>
__device__ __forceinline__ void ztile(const ZTiles &z, int t, int &opi, int &rbase)
{
    if constexpr (!GR) {
        opi = t / z.tpo[0];
        rbase = (t % z.tpo[0]) * 16;
        return;
    }
    int g = 0;
#pragma unroll
    for (int k = 1; k < kZGroups; ++k) g = (k < z.ng && t >= z.t0[k]) ? k : g;
    const int rel = t - z.t0[g];
    opi = z.op0[g] + rel / z.tpo[g];
    rbase = (rel % z.tpo[g]) * 16;
}

template <typename Tin, typename Tout, int HSP, bool GR>
__global__ void __launch_bounds__(64 * kMfma2Waves) k_zdense_mfma2(const DenseOp *__restrict__ ops, int ntiles,
                                                                   ZTiles zt)
{
    constexpr int LDM = Mfma2<HSP>::LDM, XS = Mfma2<HSP>::XS;
    __shared__ double Ms[HSP * LDM];
    __shared__ double Ma[HSP * LDM];
    __shared__ __attribute__((aligned(16))) double X[kMfma2Waves * 8 * XS];
    mfma2_stage_ops<HSP>(ops[0], Ms, Ma, threadIdx.x, 64 * kMfma2Waves);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: tile indices stay scalar
    double *Xw = X + wave * 8 * XS;
    for (int t = blockIdx.x * kMfma2Waves + wave; t < ntiles; t += gridDim.x * kMfma2Waves) {
        int opi, rbase;
        ztile<GR>(zt, t, opi, rbase);
        const DenseOp &op = ops[opi];
        mfma2_tile<Tin, Tout, HSP>(op, rbase, static_cast<Tout *>(op.dst), Ms, Ma, Xw, lane);
    }
}

// The same with the next tile's loads in flight under each tile's products
// (issued once the tile is staged, into the registers the staging has just
// freed; past the launch's end the last tile is loaded again, so no branch
// makes the compiler wait for the loads early).  FP32 rows of n = N points
// (N a multiple of 16, at most 2 HSP) in whole 16-row tiles and the FP64
// factor only (the host checks).
template <typename Tin, int HSP, int N, bool GR>
__global__ void __launch_bounds__(64 * kMfma2Waves) k_zdense_mfma2p(const DenseOp *__restrict__ ops, int ntiles,
                                                                    ZTiles zt)
{
    constexpr int LDM = Mfma2<HSP>::LDM, XS = Mfma2<HSP>::XS;
    __shared__ double Ms[HSP * LDM];
    __shared__ double Ma[HSP * LDM];
    __shared__ __attribute__((aligned(16))) double X[kMfma2Waves * 8 * XS];
    mfma2_stage_ops<HSP>(ops[0], Ms, Ma, threadIdx.x, 64 * kMfma2Waves);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *Xw = X + wave * 8 * XS;
    int t = blockIdx.x * kMfma2Waves + wave;
    if (t >= ntiles) return;   // the whole wave; no barrier follows
    // only the slice (read once, from HBM) is prefetched; the factor, reused
    // by every step of a block, is loaded at staging
    Pair<Tin> xn[16];
    auto load_x = [&](int u) {
        int oi, rb;
        ztile<GR>(zt, u, oi, rb);
        const DenseOp &o = ops[oi];
        const int n = o.n, nr = min(16, o.nrows - rb);
        const unsigned o0 = unsigned(rb) * unsigned(n) + unsigned(2 * lane < n ? 2 * lane : 0);
        const Tin *src = static_cast<const Tin *>(o.src);
#pragma unroll
        for (int r = 0; r < 16; ++r) xn[r] = ldg_pair_nt(src, o0 + unsigned(r < nr ? r : 0) * unsigned(n));
    };
    load_x(t);
    for (; t < ntiles; t += gridDim.x * kMfma2Waves) {
        int opi, rbase;
        ztile<GR>(zt, t, opi, rbase);
        const DenseOp &op = ops[opi];
        Z2Regs<Tin> R;
        {
            const int n = op.n, nr = min(16, op.nrows - rbase);
            const unsigned o0 = unsigned(rbase) * unsigned(n) + unsigned(2 * lane < n ? 2 * lane : 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                R.xa[r] = xn[r];
                R.ea[r] = ldg_pair(op.epre, o0 + unsigned(r < nr ? r : 0) * unsigned(n));
            }
        }
        double sf[HSP / 4], af[HSP / 4];
        mfma2_stage<Tin, HSP>(op, rbase, R, Xw, lane, sf, af);
        load_x(min(t + int(gridDim.x) * kMfma2Waves, ntiles - 1));
        mfma2_mma_store<float, HSP, N>(op, rbase, static_cast<float *>(op.dst), Ms, Ma, Xw, lane, sf, af);
    }
}

// ------------------------------------------- periodic FFT, one line per wave
//
// Periodic axes (x, y, and z with MASK 0) of lengths known at compile time.
// A workgroup of 8 wavefronts takes 16 consecutive real grid lines (16
// contiguous doubles per grid row for strided x lines), packs them two per
// complex line and gives each wavefront one complex line.  Stockham passes
// have compile-time radices, so all index arithmetic folds to shifts and
// multiplies; K is applied in the last forward pass, 1/L and E on store.

template <int L> struct FftPlan;
template <> struct FftPlan<16> { static constexpr int n = 2; static constexpr int R[2] = {4, 4}; };
template <> struct FftPlan<32> { static constexpr int n = 3; static constexpr int R[3] = {4, 4, 2}; };
template <> struct FftPlan<48> { static constexpr int n = 3; static constexpr int R[3] = {4, 4, 3}; };
template <> struct FftPlan<64> { static constexpr int n = 3; static constexpr int R[3] = {4, 4, 4}; };
template <> struct FftPlan<80> { static constexpr int n = 3; static constexpr int R[3] = {4, 4, 5}; };
template <> struct FftPlan<96> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 2, 3}; };
template <> struct FftPlan<128> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 4, 2}; };
template <> struct FftPlan<160> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 2, 5}; };
template <> struct FftPlan<192> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 4, 3}; };
template <> struct FftPlan<256> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 4, 4}; };
template <> struct FftPlan<320> { static constexpr int n = 4; static constexpr int R[4] = {4, 4, 4, 5}; };
template <> struct FftPlan<384> { static constexpr int n = 5; static constexpr int R[5] = {4, 4, 4, 2, 3}; };
template <> struct FftPlan<512> { static constexpr int n = 5; static constexpr int R[5] = {4, 4, 4, 4, 2}; };

constexpr int kFftWaves = 8;

template <int L, int P>
struct PassInfo {
    static constexpr int R = FftPlan<L>::R[P];
    static constexpr int ns()
    {
        int v = 1;
        for (int i = 0; i < P; ++i) v *= FftPlan<L>::R[i];
        return v;
    }
    static constexpr int Ns = ns();
    static constexpr int nb = L / R;
    static constexpr int per = (nb + 63) / 64;   // butterflies per lane
};

template <int R>
__device__ __forceinline__ void dft_small(cpx<double> (&v)[R], double sg)
{
    if constexpr (R == 2) {
        cpx<double> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        cpx<double> a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
        cpx<double> t = {v[0].x - 0.5 * a.x, v[0].y - 0.5 * a.y};
        cpx<double> u = cmul_si(b, sg * Consts<double>::s3);
        v[0] = cadd(v[0], a);
        v[1] = cadd(t, u);
        v[2] = csub(t, u);
    } else if constexpr (R == 4) {
        cpx<double> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        cpx<double> s13 = cadd(v[1], v[3]), d13 = cmul_si(csub(v[1], v[3]), sg);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
    } else {   // R == 5
        const double c1 = Consts<double>::c51, c2 = Consts<double>::c52;
        const double s1 = Consts<double>::s51, s2 = Consts<double>::s52;
        cpx<double> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cpx<double> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cpx<double> t1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        cpx<double> t2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        cpx<double> u1 = cmul_si(cpx<double>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y}, sg);
        cpx<double> u2 = cmul_si(cpx<double>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y}, sg);
        v[0] = cpx<double>{v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
        v[1] = cadd(t1, u1);
        v[4] = csub(t1, u1);
        v[2] = cadd(t2, u2);
        v[3] = csub(t2, u2);
    }
}


// One Stockham pass of the wave's line `buf` (in place: read all, sync, write all).
template <int L, int P, int SIGN, bool APPLY_K>
__device__ __forceinline__ void wave_pass(cpx<double> *buf, const cpx<double> *tw, const double *K, int lane)
{
    using PI = PassInfo<L, P>;
    constexpr int R = PI::R, Ns = PI::Ns, nb = PI::nb, per = PI::per;
    constexpr int tstep = L / (Ns * R);
    cpx<double> v[per][R];
#pragma unroll
    for (int b = 0; b < per; ++b) {
        const int j = lane + 64 * b;
        if (j < nb) {
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cpx<double> x = buf[j + r * nb];
                if (r > 0) {
                    cpx<double> t = tw[r * k * tstep];
                    if (SIGN > 0) t.y = -t.y;
                    x = cmul(x, t);
                }
                v[b][r] = x;
            }
            dft_small<R>(v[b], double(SIGN));
        }
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < per; ++b) {
        const int j = lane + 64 * b;
        if (j < nb) {
            const int k = j % Ns;
            const int od = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cpx<double> x = v[b][r];
                if (APPLY_K) {
                    const double kk = K[od + r * Ns];
                    x.x *= kk;
                    x.y *= kk;
                }
                buf[od + r * Ns] = x;
            }
        }
    }
    wave_sync();
}

template <int L, int P, int SIGN>
__device__ __forceinline__ void wave_fft(cpx<double> *buf, const cpx<double> *tw, const double *K, int lane)
{
    if constexpr (P < FftPlan<L>::n) {
        constexpr bool last_fwd = (SIGN < 0) && (P == FftPlan<L>::n - 1);
        wave_pass<L, P, SIGN, last_fwd>(buf, tw, K, lane);
        wave_fft<L, P + 1, SIGN>(buf, tw, K, lane);
    }
}

template <int L, typename Tin, typename Tout>
__global__ void __launch_bounds__(64 * kFftWaves) k_fftwave(const LineOp *__restrict__ ops)
{
    constexpr int NL = 2 * kFftWaves;   // real lines per workgroup
    const LineOp &op = ops[blockIdx.y];
    const int g0 = blockIdx.x * NL;
    if (g0 >= op.nlines) return;
    // line stride L + 1: the 16 lanes that touch one grid row hit 16 different bank groups
    constexpr int LS = L + 1;
    __shared__ cpx<double> buf[kFftWaves * LS];
    __shared__ cpx<double> tw[L];
    __shared__ double K[L];
    __shared__ int lbase[NL];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nl = min(NL, op.nlines - g0);
    const Tin *__restrict__ src = static_cast<const Tin *>(op.src);
    Tout *__restrict__ dst = static_cast<Tout *>(op.dst);
    for (int e = tid; e < L; e += 64 * kFftWaves) {
        const double *twd = reinterpret_cast<const double *>(op.tw);
        tw[e] = cpx<double>{ldg(twd, 2 * e), ldg(twd, 2 * e + 1)};
        K[e] = ldg(op.K, e);
    }
    if (tid < NL) lbase[tid] = tid < nl ? line_base(op, g0 + tid) : 0;
    __syncthreads();
    const bool contiguous = (op.stride == 1);
    const int stride = op.stride;
    // all of a thread's loads issue before the first use (lines past the
    // op's end read line 0 and are zeroed)
    constexpr int NTH = 64 * kFftWaves;
    constexpr int NPT = (NL * L + NTH - 1) / NTH;
    const int mode = !op.epre ? 0 : (!op.epre2 ? 1 : (op.epre2 == op.epre ? 2 : 3));
    int ad[NPT];
    Tin xv[NPT];
    double ev[NPT];
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int e = tid + i * NTH;
        int r, t;
        if (contiguous) { r = e / L; t = e - r * L; }
        else { t = e / NL; r = e - t * NL; }
        ad[i] = (e < NL * L && r < nl) ? lbase[r] + t * stride : 0;
    }
#pragma unroll
    for (int i = 0; i < NPT; ++i) xv[i] = ldg(src, ad[i]);
    if (mode) {
#pragma unroll
        for (int i = 0; i < NPT; ++i) ev[i] = ldg(op.epre, ad[i]);
        if (mode == 2) {
#pragma unroll
            for (int i = 0; i < NPT; ++i) ev[i] *= ev[i];
        } else if (mode == 3) {
#pragma unroll
            for (int i = 0; i < NPT; ++i) ev[i] *= ldg(op.epre2, ad[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int e = tid + i * NTH;
        if (e >= NL * L) break;
        int r, t;
        if (contiguous) { r = e / L; t = e - r * L; }
        else { t = e / NL; r = e - t * NL; }
        double v = 0.0;
        if (r < nl) v = mode ? double(xv[i]) * ev[i] : double(xv[i]);
        reinterpret_cast<double *>(&buf[(r >> 1) * LS + t])[r & 1] = v;
    }
    __syncthreads();
    cpx<double> *line = buf + wave * LS;
    wave_fft<L, 0, -1>(line, tw, K, lane);
    wave_fft<L, 0, +1>(line, tw, K, lane);
    __syncthreads();   // the store mixes lines of different waves
    constexpr double scale = 1.0 / L;
    for (int e = tid; e < NL * L; e += 64 * kFftWaves) {
        int r, t;
        if (contiguous) { r = e / L; t = e - r * L; }
        else { t = e / NL; r = e - t * NL; }
        if (r >= nl) continue;
        const int a = lbase[r] + t * stride;
        double v = reinterpret_cast<const double *>(&buf[(r >> 1) * LS + t])[r & 1] * scale;
        if (op.epost) v *= ldg(op.epost, a);
        stg(dst, a, store_round<Tout>(v, op.rkey, unsigned(a)));
    }
}


// ------------------------- periodic FFT of length L = 16 P, in registers
//
// Four-step transform, L = 16 x P (P = 8: L = 128; P = 16: L = 256).  P
// lanes hold one complex line (two real grid lines): lane p holds
// x[p + P m] (m = 0..15), does a 16-point DFT over m in registers and
// multiplies by w_L^{p k2}; one LDS transpose gives lane q the columns
// c = q + P r (r < 16/P) with their P values, whose P-point DFTs yield
// X[c + 16 k1].  K is applied in that order and the inverse runs the same
// steps backwards, so the spectrum is never reordered.  Each line makes two
// LDS passes (one transpose per direction) instead of one per radix pass.
// A wave holds G = 64/P complex lines; lanes are line-minor (lane = g + G p)
// so the lanes that load one grid row of strided (x, y) lines are adjacent
// and a quarter-wave touches few cache lines.  Waves are persistent over
// units (the 2G real lines of one wave) and the next unit's loads are in
// flight during the current unit's transforms.

constexpr int kFRWaves = 4;

template <int P>
struct FReg {
    static constexpr int L = 16 * P;
    static constexpr int G = 64 / P;           // complex lines per wave
    static constexpr int R = 16 / P;           // columns per lane after the transpose
    static constexpr int RS = P + 1;           // tile row stride (complex)
    // line tiles G16 = 16/G slots (of 16 B) apart in the bank rows: the G
    // lines x 16/G points of one 16-lane b128 phase hit disjoint banks
    static constexpr int RG = 16 * RS + 16 / G;
    static constexpr int UNIT = 2 * G;         // real lines per wave unit
};

// w_16^{m} = exp(SIGN 2 pi i m / 16)
template <int SIGN>
__device__ __forceinline__ cpx<double> w16(int m)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, r2 = 0.70710678118654752440;
    const double C[16] = {1.0, c1, r2, s1, 0.0, -s1, -r2, -c1, -1.0, -c1, -r2, -s1, 0.0, s1, r2, c1};
    const double S[16] = {0.0, s1, r2, c1, 1.0, c1, r2, s1, 0.0, -s1, -r2, -c1, -1.0, -c1, -r2, -s1};
    return cpx<double>{C[m & 15], SIGN * S[m & 15]};
}

// In-place 16-point DFT of v[o .. o+15] with w = exp(SIGN 2 pi i / 16).
template <int SIGN, int N>
__device__ __forceinline__ void dft16(cpx<double> (&v)[N], int o = 0)
{
    cpx<double> t[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        cpx<double> a[4] = {v[o + n1], v[o + n1 + 4], v[o + n1 + 8], v[o + n1 + 12]};
        dft_small<4>(a, double(SIGN));
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) t[n1][k2] = (n1 * k2 == 0) ? a[k2] : cmul(a[k2], w16<SIGN>(n1 * k2));
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        cpx<double> a[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
        dft_small<4>(a, double(SIGN));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) v[o + k2 + 4 * k1] = a[k1];
    }
}

// In-place 8-point DFT of v[o .. o+7]: radix 4 over n2 (n = n1 + 2 n2), w_8
// twiddles, radix 2 over n1.
template <int SIGN, int N>
__device__ __forceinline__ void dft8(cpx<double> (&v)[N], int o)
{
    cpx<double> t[2][4];
#pragma unroll
    for (int n1 = 0; n1 < 2; ++n1) {
        cpx<double> a[4] = {v[o + n1], v[o + n1 + 2], v[o + n1 + 4], v[o + n1 + 6]};
        dft_small<4>(a, double(SIGN));
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) t[n1][k2] = (n1 * k2 == 0) ? a[k2] : cmul(a[k2], w16<SIGN>(2 * n1 * k2));
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        v[o + k2] = cadd(t[0][k2], t[1][k2]);
        v[o + k2 + 4] = csub(t[0][k2], t[1][k2]);
    }
}

template <int P, int SIGN, int N>
__device__ __forceinline__ void dft_p(cpx<double> (&v)[N], int o)
{
    if constexpr (P == 16) dft16<SIGN>(v, o);
    else dft8<SIGN>(v, o);
}

// Loads of one unit held until the unit starts.
template <typename Tin>
struct FRegLoads {
    Pair<Tin> x[16];
};

// Unit u of the launch: op u / upo, real lines UNIT (u % upo) + 2 g, +1.
template <int P, typename Tin>
__device__ __forceinline__ void freg_fetch(const LineOp *__restrict__ ops, int upo, int u, int nunits, int g, int p,
                                           FRegLoads<Tin> &R)
{
    u = min(u, nunits - 1);
    const LineOp &op = ops[u / upo];
    const int r0 = (u % upo) * FReg<P>::UNIT + 2 * g;
    const bool ok0 = r0 < op.nlines, ok1 = r0 + 1 < op.nlines;
    const int b0 = line_base(op, ok0 ? r0 : 0), b1 = line_base(op, ok1 ? r0 + 1 : 0);
    const Tin *src = static_cast<const Tin *>(op.src);
    const unsigned o0 = unsigned(b0) + unsigned(p) * unsigned(op.stride), step = unsigned(P) * unsigned(op.stride);
    if (op.pairs) {
#pragma unroll
        for (int m = 0; m < 16; ++m) R.x[m] = ldg_pair(src, o0 + m * step);
    } else {
        const unsigned o1 = o0 + unsigned(b1 - b0);
#pragma unroll
        for (int m = 0; m < 16; ++m) R.x[m] = Pair<Tin>{ldg(src, o0 + m * step), ldg(src, o1 + m * step)};
    }
}

// The twiddles w_L^{-e} of the op's length into LDS at e + e/16.
template <int P>
__device__ __forceinline__ void freg_stage_twiddles(const LineOp &op, cpx<double> *TW, int tid, int nthreads)
{
    constexpr int L = FReg<P>::L;
    const double *twd = reinterpret_cast<const double *>(op.tw);
    for (int e = tid; e < L; e += nthreads) TW[e + (e >> 4)] = cpx<double>{ldg(twd, 2 * e), ldg(twd, 2 * e + 1)};
}

// The line operator of one unit: lane (g, p) holds x[p + P m] of its
// complex line (real lines r0, r0 + 1) in v; the result goes to op.dst.
template <int P, typename Tout>
__device__ __forceinline__ void freg_apply(const LineOp &op, int r0, cpx<double> (&v)[16], cpx<double> *Tg,
                                           const cpx<double> *TW, int p)
{
    using F = FReg<P>;
    constexpr int L = F::L, R = F::R, RS = F::RS;
    const bool ok0 = r0 < op.nlines, ok1 = r0 + 1 < op.nlines;
    const int b0 = line_base(op, ok0 ? r0 : 0), b1 = line_base(op, ok1 ? r0 + 1 : 0);
    const int stride = op.stride;
    const unsigned o0 = unsigned(b0) + unsigned(p) * unsigned(stride), step = unsigned(P) * unsigned(stride);
    const unsigned o1 = o0 + unsigned(b1 - b0);
    if (op.epre) {   // first axis of a step (periodic z): not prefetched
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            v[m].x *= pre_factor(op.epre, op.epre2, o0 + m * step);
            v[m].y *= pre_factor(op.epre, op.epre2, o1 + m * step);
        }
    }
    // forward: over m, twiddle w^{-p k2}, transpose, over p (per column)
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        const int e = (p * k) & (L - 1);
        v[k] = cmul(v[k], TW[e + (e >> 4)]);
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) Tg[k * RS + p] = v[k];
    wave_sync();
    const int q = p;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < P; ++k) v[r * P + k] = Tg[(q + P * r) * RS + k];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dft_p<P, -1>(v, r * P);   // v[r P + k1] = X[c + 16 k1], c = q + P r
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            const double kk = ldg(op.K, unsigned(q + P * r + 16 * k1)) * (1.0 / L);   // 1/L folded in (exact)
            v[r * P + k1].x *= kk;
            v[r * P + k1].y *= kk;
        }
        dft_p<P, +1>(v, r * P);   // over k1, output index p
    }
    // inverse: twiddle w^{+p c}, transpose back, over k2
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 1; k < P; ++k) {
            const int e = (k * (q + P * r)) & (L - 1);
            cpx<double> w = TW[e + (e >> 4)];
            w.y = -w.y;
            v[r * P + k] = cmul(v[r * P + k], w);
        }
    wave_sync();
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int k = 0; k < P; ++k) Tg[(q + P * r) * RS + k] = v[r * P + k];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = Tg[k * RS + p];
    dft16<+1>(v);   // v[m] = x[p + P m] after the operator
    Tout *dst = static_cast<Tout *>(op.dst);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        double y0 = v[m].x, y1 = v[m].y;
        if (op.epost) {
            y0 *= ldg(op.epost, o0 + m * step);
            y1 *= ldg(op.epost, o1 + m * step);
        }
        const unsigned a0 = o0 + m * step, a1 = o1 + m * step;
        if (op.pairs) {
            if (ok0) stg_pair(dst, a0, store_round<Tout>(y0, op.rkey, a0), store_round<Tout>(y1, op.rkey, a0 + 1));
        } else {
            if (ok0) stg(dst, a0, store_round<Tout>(y0, op.rkey, a0));
            if (ok1) stg(dst, a1, store_round<Tout>(y1, op.rkey, a1));
        }
    }
}

template <int P, typename Tin, typename Tout>
__global__ void __launch_bounds__(64 * kFRWaves) k_fftreg(const LineOp *__restrict__ ops, int nunits, int upo)
{
    using F = FReg<P>;
    constexpr int L = F::L, G = F::G, RG = F::RG;
    __shared__ cpx<double> T[kFRWaves * G * RG];
    __shared__ cpx<double> TW[L + L / 16];   // w_L^{-e} at e + e/16 (one pad per 16 entries)
    const int tid = threadIdx.x, lane = tid & 63;
    freg_stage_twiddles<P>(ops[0], TW, tid, 64 * kFRWaves);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane % G, p = lane / G;   // line-minor
    cpx<double> *Tg = T + (wave * G + g) * RG;
    const int ustride = gridDim.x * kFRWaves;
    int u = blockIdx.x * kFRWaves + wave;
    if (u >= nunits) return;   // no workgroup barrier follows
    FRegLoads<Tin> ld;
    freg_fetch<P, Tin>(ops, upo, u, nunits, g, p, ld);
    for (; u < nunits; u += ustride) {
        const LineOp &op = ops[u / upo];
        const int r0 = (u % upo) * F::UNIT + 2 * g;
        cpx<double> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = cpx<double>{double(ld.x[m].x), double(ld.x[m].y)};
        freg_fetch<P, Tin>(ops, upo, u + ustride, nunits, g, p, ld);   // in flight during the transforms
        freg_apply<P, Tout>(op, r0, v, Tg, TW, p);
    }
}

// ------- register FFT, lengths P B (16 x 12, 16 x 20, 8 x 12, 8 x 8: 192, 320, 96, 64)
//
// Periodic lines of length L = P B: P lanes per complex line, lane p
// holding x[p + P m] for m < B.  A B-point DFT per lane (radix 4 x B/4, or
// the 8-point form), the twiddle w_L^{p k2}, one LDS transpose, then per
// column k2 a P-point DFT over p, K at spectrum index k2 + B k1, and the
// inverse P-point DFT; the transpose back, the inverse twiddle and the
// inverse B-point DFT.  Column k2 = q + P r goes to lane q: when B > P
// lanes q < B - P take a second column.

// w_N^m = exp(SIGN 2 pi i m / N), N = 12 or 20
template <int N, int SIGN>
__device__ __forceinline__ cpx<double> wN(int m)
{
    if constexpr (N == 12) {
        constexpr double a = 0.8660254037844386467637;
        const double C[12] = {1.0, a, 0.5, 0.0, -0.5, -a, -1.0, -a, -0.5, 0.0, 0.5, a};
        const double S[12] = {0.0, 0.5, a, 1.0, a, 0.5, 0.0, -0.5, -a, -1.0, -a, -0.5};
        return cpx<double>{C[m % 12], SIGN * S[m % 12]};
    } else {
        constexpr double a = 0.9510565162951535721164, b = 0.8090169943749474241023;
        constexpr double c = 0.5877852522924731291687, d = 0.3090169943749474241023;
        const double C[20] = {1.0, a, b, c, d, 0.0, -d, -c, -b, -a, -1.0, -a, -b, -c, -d, 0.0, d, c, b, a};
        const double S[20] = {0.0, d, c, b, a, 1.0, a, b, c, d, 0.0, -d, -c, -b, -a, -1.0, -a, -b, -c, -d};
        return cpx<double>{C[m % 20], SIGN * S[m % 20]};
    }
}

// In-place N-point DFT of v[0 .. N-1], N = 4 N2 (N2 = 3, 5): n = n1 + 4 n2
// (DFTs of N2 over n2), twiddle w_N^{n1 k2}, DFTs of 4 over n1; output
// index k = k2 + N2 k1.
template <int N, int SIGN>
__device__ __forceinline__ void dft_4x(cpx<double> (&v)[N])
{
    constexpr int N2 = N / 4;
    cpx<double> t[4][N2];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        cpx<double> a[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) a[n2] = v[n1 + 4 * n2];
        dft_small<N2>(a, double(SIGN));
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) t[n1][k2] = (n1 * k2 == 0) ? a[k2] : cmul(a[k2], wN<N, SIGN>(n1 * k2));
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
        cpx<double> b[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
        dft_small<4>(b, double(SIGN));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) v[k2 + N2 * k1] = b[k1];
    }
}

// The B-point DFT of one lane: radix 4 x B/4, or the 8-point form.
template <int B, int SIGN>
__device__ __forceinline__ void dft_lane(cpx<double> (&v)[B])
{
    if constexpr (B == 8) dft8<SIGN>(v, 0);
    else dft_4x<B, SIGN>(v);
}

template <int B, int P = 16>
struct FReg2 {
    static constexpr int L = P * B;
    static constexpr int G = 64 / P;            // complex lines per wave (P lanes each)
    static constexpr int RS = P + 1;            // tile row stride (complex)
    static constexpr int RG = B * RS + 16 / G;  // line tiles 16/G slots apart in the bank rows
    static constexpr int UNIT = 2 * G;          // real lines per wave unit
    static constexpr int R = (B + P - 1) / P;   // columns per lane
    static constexpr int W = B > 16 ? 3 : 4;    // waves per workgroup: two workgroups per CU in LDS
    static constexpr int WG_PER_CU = L <= 64 ? 4 : 2;   // workgroups that fit one CU's LDS
};

template <int B, typename Tin>
struct FReg2Loads {
    Pair<Tin> x[B];
};

template <int B, int P, typename Tin>
__device__ __forceinline__ void freg2_fetch(const LineOp *__restrict__ ops, int upo, int u, int nunits, int g, int p,
                                            FReg2Loads<B, Tin> &R)
{
    u = min(u, nunits - 1);
    const LineOp &op = ops[u / upo];
    const int r0 = (u % upo) * FReg2<B, P>::UNIT + 2 * g;
    const bool ok0 = r0 < op.nlines, ok1 = r0 + 1 < op.nlines;
    const int b0 = line_base(op, ok0 ? r0 : 0), b1 = line_base(op, ok1 ? r0 + 1 : 0);
    const Tin *src = static_cast<const Tin *>(op.src);
    const unsigned o0 = unsigned(b0) + unsigned(p) * unsigned(op.stride), step = unsigned(P) * unsigned(op.stride);
    if (op.pairs) {
#pragma unroll
        for (int m = 0; m < B; ++m) R.x[m] = ldg_pair(src, o0 + m * step);
    } else {
        const unsigned o1 = o0 + unsigned(b1 - b0);
#pragma unroll
        for (int m = 0; m < B; ++m) R.x[m] = Pair<Tin>{ldg(src, o0 + m * step), ldg(src, o1 + m * step)};
    }
}

// The twiddles w_L^{-e} of an op's length-L table into LDS at e + e/16.
template <int L>
__device__ __forceinline__ void stage_twiddles(const LineOp &op, cpx<double> *TW, int tid, int nthreads)
{
    const double *twd = reinterpret_cast<const double *>(op.tw);
    for (int e = tid; e < L; e += nthreads) TW[e + (e >> 4)] = cpx<double>{ldg(twd, 2 * e), ldg(twd, 2 * e + 1)};
}

template <int B, int P, typename Tout>
__device__ __forceinline__ void freg2_apply(const LineOp &op, int r0, cpx<double> (&v)[B], cpx<double> *Tg,
                                            const cpx<double> *TW, int p)
{
    using F = FReg2<B, P>;
    constexpr int L = F::L, RS = F::RS, R = F::R;
    const bool ok0 = r0 < op.nlines, ok1 = r0 + 1 < op.nlines;
    const int b0 = line_base(op, ok0 ? r0 : 0), b1 = line_base(op, ok1 ? r0 + 1 : 0);
    const unsigned o0 = unsigned(b0) + unsigned(p) * unsigned(op.stride), step = unsigned(P) * unsigned(op.stride);
    const unsigned o1 = o0 + unsigned(b1 - b0);
    if (op.epre) {   // first axis of a step
#pragma unroll
        for (int m = 0; m < B; ++m) {
            v[m].x *= pre_factor(op.epre, op.epre2, o0 + m * step);
            v[m].y *= pre_factor(op.epre, op.epre2, o1 + m * step);
        }
    }
    dft_lane<B, -1>(v);   // over m: v[k2]
#pragma unroll
    for (int k = 1; k < B; ++k) {
        const int e = p * k;   // < L
        v[k] = cmul(v[k], TW[e + (e >> 4)]);
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < B; ++k) Tg[k * RS + p] = v[k];
    wave_sync();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int c = p + P * r;   // column k2 of this lane
        if (c < B) {
            cpx<double> col[P];
#pragma unroll
            for (int k = 0; k < P; ++k) col[k] = Tg[c * RS + k];
            dft_p<P, -1>(col, 0);   // over p: spectrum index c + B k1
#pragma unroll
            for (int k1 = 0; k1 < P; ++k1) {
                const double kk = ldg(op.K, unsigned(c + B * k1)) * (1.0 / L);   // 1/L folded in
                col[k1].x *= kk;
                col[k1].y *= kk;
            }
            dft_p<P, +1>(col, 0);
#pragma unroll
            for (int k = 0; k < P; ++k) Tg[c * RS + k] = col[k];
        }
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < B; ++k) v[k] = Tg[k * RS + p];
#pragma unroll
    for (int k = 1; k < B; ++k) {
        const int e = p * k;
        cpx<double> w = TW[e + (e >> 4)];
        w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
    dft_lane<B, +1>(v);   // v[m] = x[p + P m] after the operator
    Tout *dst = static_cast<Tout *>(op.dst);
#pragma unroll
    for (int m = 0; m < B; ++m) {
        double y0 = v[m].x, y1 = v[m].y;
        if (op.epost) {
            y0 *= ldg(op.epost, o0 + m * step);
            y1 *= ldg(op.epost, o1 + m * step);
        }
        const unsigned a0 = o0 + m * step, a1 = o1 + m * step;
        if (op.pairs) {
            if (ok0) stg_pair(dst, a0, store_round<Tout>(y0, op.rkey, a0), store_round<Tout>(y1, op.rkey, a0 + 1));
        } else {
            if (ok0) stg(dst, a0, store_round<Tout>(y0, op.rkey, a0));
            if (ok1) stg(dst, a1, store_round<Tout>(y1, op.rkey, a1));
        }
    }
}

template <int B, int P, typename Tin, typename Tout>
__global__ void __launch_bounds__((64 * FReg2<B, P>::W)) k_fftreg2(const LineOp *__restrict__ ops, int nunits, int upo)
{
    using F = FReg2<B, P>;
    constexpr int L = F::L, G = F::G, RG = F::RG, W = F::W;
    __shared__ cpx<double> T[W * G * RG];
    __shared__ cpx<double> TW[L + L / 16];
    const int tid = threadIdx.x, lane = tid & 63;
    stage_twiddles<L>(ops[0], TW, tid, 64 * W);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane % G, p = lane / G;   // line-minor
    cpx<double> *Tg = T + (wave * G + g) * RG;
    const int ustride = gridDim.x * W;
    int u = blockIdx.x * W + wave;
    if (u >= nunits) return;   // no workgroup barrier follows
    FReg2Loads<B, Tin> ld;
    freg2_fetch<B, P, Tin>(ops, upo, u, nunits, g, p, ld);
    for (; u < nunits; u += ustride) {
        const LineOp &op = ops[u / upo];
        const int r0 = (u % upo) * F::UNIT + 2 * g;
        cpx<double> v[B];
#pragma unroll
        for (int m = 0; m < B; ++m) v[m] = cpx<double>{double(ld.x[m].x), double(ld.x[m].y)};
        freg2_fetch<B, P, Tin>(ops, upo, u + ustride, nunits, g, p, ld);   // in flight during the transforms
        freg2_apply<B, P, Tout>(op, r0, v, Tg, TW, p);
    }
}

// ------------------------- one launch for the x lines of lengths 192, 256, 320
//
// A batch of uneven NX (the phase-diagram sweep: NX 192, 256 and 320) has
// one periodic last-axis transform per length in each contour step.
// k_fftmix runs them as one launch: workgroups [wg0, wg0 + nwg) of group k
// run the persistent-wave loop of that length's register FFT (the bodies of
// k_fftreg<16> and k_fftreg2<12,16> / <20,16>) over the group's ops, the
// workgroups split between the groups in proportion to their work per
// wave.  Each launch alone would take every CU and end in a partial round
// of units; side by side on streams they overlapped only in part.  LDS is
// the largest of the three layouts (the 256-point tiles), so two
// workgroups share a CU as in each kernel alone; the 320-point group runs
// three of the four waves, as k_fftreg2<20,16> does.

struct MixGroup {
    int op0;        // first op of the group, from the launch's first op
    int nunits;     // wave units of the group
    int upo;        // units per op
    int wg0, nwg;   // the group's workgroups
};

struct MixArgs {
    MixGroup g[3];  // lengths 256, 192, 320 (nwg 0: absent)
};

constexpr int kMixThreads = 64 * kFRWaves;

template <int... N>
constexpr int max_of() { int m = 0; ((m = N > m ? N : m), ...); return m; }

constexpr int kMixLds = max_of<int((kFRWaves * FReg<16>::G * FReg<16>::RG + FReg<16>::L + FReg<16>::L / 16) *
                                   sizeof(cpx<double>)),
                               int((FReg2<12>::W * FReg2<12>::G * FReg2<12>::RG + FReg2<12>::L + FReg2<12>::L / 16) *
                                   sizeof(cpx<double>)),
                               int((FReg2<20>::W * FReg2<20>::G * FReg2<20>::RG + FReg2<20>::L + FReg2<20>::L / 16) *
                                   sizeof(cpx<double>))>();
static_assert(FReg2<12>::W <= kFRWaves && FReg2<20>::W <= kFRWaves, "mixed launch has four waves per workgroup");

// k_fftreg<16>'s loop over the units of workgroup vb of nvb.
template <typename Tin, typename Tout>
__device__ __forceinline__ void mix_run256(const LineOp *__restrict__ ops, int nunits, int upo, int vb, int nvb,
                                           cpx<double> *smem)
{
    using F = FReg<16>;
    constexpr int G = F::G, RG = F::RG;
    cpx<double> *T = smem, *TW = smem + kFRWaves * G * RG;
    const int tid = threadIdx.x, lane = tid & 63;
    freg_stage_twiddles<16>(ops[0], TW, tid, kMixThreads);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane % G, p = lane / G;
    cpx<double> *Tg = T + (wave * G + g) * RG;
    const int ustride = nvb * kFRWaves;
    int u = vb * kFRWaves + wave;
    if (u >= nunits) return;   // no workgroup barrier follows
    FRegLoads<Tin> ld;
    freg_fetch<16, Tin>(ops, upo, u, nunits, g, p, ld);
    for (; u < nunits; u += ustride) {
        const LineOp &op = ops[u / upo];
        const int r0 = (u % upo) * F::UNIT + 2 * g;
        cpx<double> v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = cpx<double>{double(ld.x[m].x), double(ld.x[m].y)};
        freg_fetch<16, Tin>(ops, upo, u + ustride, nunits, g, p, ld);
        freg_apply<16, Tout>(op, r0, v, Tg, TW, p);
    }
}

// k_fftreg2<B, 16>'s loop over the units of workgroup vb of nvb (waves
// W and up idle after the twiddle barrier).
template <int B, typename Tin, typename Tout>
__device__ __forceinline__ void mix_run2(const LineOp *__restrict__ ops, int nunits, int upo, int vb, int nvb,
                                         cpx<double> *smem)
{
    using F = FReg2<B, 16>;
    constexpr int L = F::L, G = F::G, RG = F::RG, W = F::W;
    cpx<double> *T = smem, *TW = smem + W * G * RG;
    const int tid = threadIdx.x, lane = tid & 63;
    stage_twiddles<L>(ops[0], TW, tid, kMixThreads);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (wave >= W) return;   // no workgroup barrier follows
    const int g = lane % G, p = lane / G;
    cpx<double> *Tg = T + (wave * G + g) * RG;
    const int ustride = nvb * W;
    int u = vb * W + wave;
    if (u >= nunits) return;
    FReg2Loads<B, Tin> ld;
    freg2_fetch<B, 16, Tin>(ops, upo, u, nunits, g, p, ld);
    for (; u < nunits; u += ustride) {
        const LineOp &op = ops[u / upo];
        const int r0 = (u % upo) * F::UNIT + 2 * g;
        cpx<double> v[B];
#pragma unroll
        for (int m = 0; m < B; ++m) v[m] = cpx<double>{double(ld.x[m].x), double(ld.x[m].y)};
        freg2_fetch<B, 16, Tin>(ops, upo, u + ustride, nunits, g, p, ld);
        freg2_apply<B, 16, Tout>(op, r0, v, Tg, TW, p);
    }
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kMixThreads) k_fftmix(const LineOp *__restrict__ ops, MixArgs a)
{
    __shared__ cpx<double> smem[kMixLds / sizeof(cpx<double>)];
    const int b = blockIdx.x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const MixGroup &gr = a.g[k];
        if (b < gr.wg0 || b >= gr.wg0 + gr.nwg) continue;
        if (k == 0) mix_run256<Tin, Tout>(ops + gr.op0, gr.nunits, gr.upo, b - gr.wg0, gr.nwg, smem);
        else if (k == 1) mix_run2<12, Tin, Tout>(ops + gr.op0, gr.nunits, gr.upo, b - gr.wg0, gr.nwg, smem);
        else mix_run2<20, Tin, Tout>(ops + gr.op0, gr.nunits, gr.upo, b - gr.wg0, gr.nwg, smem);
        return;
    }
}

// ------------------------------------------- fused step on a periodic 64 x 64 plane
//
// One contour step of one chain on a periodic 2D plane (NY = 1) of 64 x 64
// points in one workgroup: the slice times its pre-factor into LDS
// (coalesced), the 64-point z transforms with Kz over the rows (two real
// rows per complex line), the 64-point x transforms with Kx over the
// columns (two real columns per complex line), and the plane out
// (coalesced).  Each complex line runs the 8 x 8 register FFT of k_fftreg2
// with its transpose done in the line's own LDS cells.  It replaces the z
// and x launches of a step: the scratch slice between the two axes stays
// on chip.

struct PlaneOp {
    const void *src;           // slice t (T), x-major, z fastest
    void *dst;                 // slice t + 1 (T)
    const double *epre;        // E of this step
    const double *epre2;       // factor the stored slice owes (null, epre, or another E)
    const double *Kz, *Kx;     // 64-point spectra, natural order
    const double *twz, *twx;   // w_64^{-e}, complex interleaved
    unsigned rkey;             // unbiased-rounding key of FP32 stores (store_round)
};

constexpr int kPlaneN = 64;        // points per axis
constexpr int kPlanePS = 66;       // LDS row stride (doubles): the rows of a half-wave start 8 banks apart
constexpr int kPlaneWaves = 4;     // 32 complex lines per axis, 8 per wave

// The operator on one complex line whose slot j (0..63) lives at Pl[a + j st]
// (real part) and Pl[b + j st] (imaginary part).  Lane p holds slots p + 8 m
// in v on entry and the operated values there on exit; the transpose uses
// the line's own cells (each lane writes the cells it read).
__device__ __forceinline__ void plane_line(cpx<double> (&v)[8], double *Pl, int a, int b, int st,
                                           const cpx<double> *TW, const double *__restrict__ K, int p)
{
    dft8<-1>(v, 0);   // over m: v[k2]
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        const int e = p * k;
        v[k] = cmul(v[k], TW[e + (e >> 4)]);
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = (k * 8 + p) * st;
        Pl[a + j] = v[k].x;
        Pl[b + j] = v[k].y;
    }
    wave_sync();
    cpx<double> col[8];   // column k2 = p: spectrum index p + 8 k1
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = (p * 8 + k) * st;
        col[k] = cpx<double>{Pl[a + j], Pl[b + j]};
    }
    dft8<-1>(col, 0);
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
        const double kk = ldg(K, unsigned(p + 8 * k1)) * (1.0 / 64);   // 1/L folded in (exact)
        col[k1].x *= kk;
        col[k1].y *= kk;
    }
    dft8<+1>(col, 0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = (p * 8 + k) * st;
        Pl[a + j] = col[k].x;
        Pl[b + j] = col[k].y;
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int j = (k * 8 + p) * st;
        v[k] = cpx<double>{Pl[a + j], Pl[b + j]};
    }
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        const int e = p * k;
        cpx<double> w = TW[e + (e >> 4)];
        w.y = -w.y;
        v[k] = cmul(v[k], w);
    }
    dft8<+1>(v, 0);   // v[m] = slot p + 8 m after the operator
}

// grid: persistent workgroups over the ops of the launch
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(64 * kPlaneWaves) k_plane64(const PlaneOp *__restrict__ ops, int nops)
{
    constexpr int N = kPlaneN, PS = kPlanePS, NT = 64 * kPlaneWaves;
    __shared__ double Pl[N * PS];
    __shared__ cpx<double> TWz[N + N / 16], TWx[N + N / 16];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane & 7, p = lane >> 3;   // line-minor
    for (int e = tid; e < N; e += NT) {      // every op of a launch shares the tables
        TWz[e + (e >> 4)] = cpx<double>{ldg(ops[0].twz, 2 * e), ldg(ops[0].twz, 2 * e + 1)};
        TWx[e + (e >> 4)] = cpx<double>{ldg(ops[0].twx, 2 * e), ldg(ops[0].twx, 2 * e + 1)};
    }
    const int c = wave * 8 + g;   // complex line of this lane: rows / columns 2c, 2c + 1
    for (int o = blockIdx.x; o < nops; o += gridDim.x) {
        const PlaneOp &op = ops[o];
        const Tin *src = static_cast<const Tin *>(op.src);
        for (int i = tid; i < N * N; i += NT)
            Pl[(i >> 6) * PS + (i & 63)] = double(ldg(src, unsigned(i))) * pre_factor(op.epre, op.epre2, unsigned(i));
        __syncthreads();
        {   // z: rows x = 2c, 2c + 1
            const int a = 2 * c * PS, b = a + PS;
            cpx<double> v[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) v[m] = cpx<double>{Pl[a + p + 8 * m], Pl[b + p + 8 * m]};
            plane_line(v, Pl, a, b, 1, TWz, op.Kz, p);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                Pl[a + p + 8 * m] = v[m].x;
                Pl[b + p + 8 * m] = v[m].y;
            }
        }
        __syncthreads();
        {   // x: columns z = 2c, 2c + 1
            const int a = 2 * c, b = a + 1;
            cpx<double> v[8];
#pragma unroll
            for (int m = 0; m < 8; ++m) v[m] = cpx<double>{Pl[a + (p + 8 * m) * PS], Pl[b + (p + 8 * m) * PS]};
            plane_line(v, Pl, a, b, PS, TWx, op.Kx, p);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                Pl[a + (p + 8 * m) * PS] = v[m].x;
                Pl[b + (p + 8 * m) * PS] = v[m].y;
            }
        }
        __syncthreads();
        Tout *dst = static_cast<Tout *>(op.dst);
        for (int i = tid; i < N * N; i += NT)
            stg(dst, unsigned(i), store_round<Tout>(Pl[(i >> 6) * PS + (i & 63)], op.rkey, unsigned(i)));
        __syncthreads();   // the next op overwrites the plane
    }
}

// ------------------------------- fused z + x step on 2D films, split by z columns
//
// One contour step of one chain on a 2D plane (NY = 1) of NX = L rows (x,
// periodic, L = 192, 256 or 320) by NZ = 128 points (z, the dense operator
// of k_zdense_mfma2: symmetric and antisymmetric halves of 64 x 64).  The
// plane's step is split by z columns into kFilmParts workgroups: part
// k0 = 16 part computes the z operator's symmetric output columns
// k0..k0+15, i.e. the z points k0..k0+15 and n-16-k0..n-1-k0 (32 columns),
// for every row, keeps them in LDS, then runs the x transforms of those 32
// columns (16 complex lines) and stores them: the scratch between the two
// axes never leaves the CU.  Every part reads the whole input plane (each
// z output needs whole rows), so the parts of one plane go to one XCD
// (block ids o % 8 + 8 part + 32 (o / 8): the dispatcher deals consecutive
// workgroups to the eight XCDs in turn) and share the reads through its L2.
// HBM then moves the input slice, the factor and the output slice once: 16
// bytes per voxel-step in FP32-split, against 24 with the scratch slice
// written by the z launch and read back by the x launch.
//
// The arithmetic is that of the two launches (k_zdense_mfma2 followed by
// k_fftreg<16> for L = 256 or k_fftreg2<L/16, 16> otherwise): the same MFMA
// operands in the same order, the scratch rounded to T, the same transform
// steps and the same stores with the same rounding keys, so the results are
// bit for bit those of the split step.

constexpr int kFilmParts = 4;    // workgroups per plane and step
constexpr int kFilmWaves = 4;
constexpr int kFilmN = 128;      // NZ
constexpr int kFilmCols = 32;    // z columns of one part (LDS row length, in T)

struct FilmOp {
    const void *src;       // slice t (T), x-major, z fastest
    void *dst;             // slice t + 1 (T)
    const double *epre;    // E of this step
    const double *epre2;   // factor the stored slice owes (null, epre, or another E)
    const float *e2f;      // E^2 in FP32 (inside a block, FP32-split with the FP32 factor on), or null
    const double *Kx;      // x spectrum of this slot (length L)
    unsigned rkey;         // unbiased-rounding key of the stored slice (store_round)
};

struct FilmTab {           // shared by every op of a launch
    const double *MsT, *MaT;   // z operator halves, [64][ldm], MsT[j][k] = Ms[k][j]
    const double *tw;          // w_L^{-e}, complex interleaved
    int ldm;
};

// Slot of complex line i (local z columns 2i, 2i+1) in LDS row e.  The 16
// slots of a row are permuted per row so that the rows one LDS access phase
// touches (4 rows of 16-byte cells, or 8 rows of 8-byte cells) start in
// different banks.
template <typename T>
__device__ __forceinline__ int film_slot(int e, int i)
{
    return i ^ (4 * ((sizeof(T) == 8 ? e : e >> 1) & 3));
}

// LDS offset (in T) of local column c of row e.
template <typename T>
__device__ __forceinline__ int film_at(int e, int c)
{
    return e * kFilmCols + 2 * film_slot<T>(e, c >> 1) + (c & 1);
}

// The z operator on rows rbase..rbase+15 for this part's 16 symmetric
// output columns (mfma2_tile's staging and fragments; bs/ba hold the part's
// column of Ms^T / Ma^T for each K step); results rounded to T into the
// part's LDS rows.  The slice is loaded as ordinary (cached) data: the
// other parts of the plane read the same rows.  film_zload loads a tile's
// slice rows (the caller keeps the next tile's in flight), film_zstage
// applies the step's factor (loaded there: every step of a block reuses
// it) and builds the fragments, film_zmma runs the products.
template <typename T>
__device__ __forceinline__ void film_zload(const FilmOp &op, int rbase, int lane, Pair<T> (&xa)[16])
{
    const T *src = static_cast<const T *>(op.src);
    const unsigned o0 = unsigned(rbase) * unsigned(kFilmN) + unsigned(2 * lane);
#pragma unroll
    for (int r = 0; r < 16; ++r) xa[r] = ldg_pair(src, o0 + unsigned(r) * unsigned(kFilmN));
}

template <typename T>
__device__ __forceinline__ void film_zstage(const FilmOp &op, int rbase, const Pair<T> (&xa)[16], double *Xw,
                                            int lane, double (&sf)[16], double (&af)[16])
{
    constexpr int n = kFilmN, NT = 16, XS = Mfma2<64>::XS;
    const int li = lane & 15, lj = lane >> 4;
    const int c = 2 * lane;   // this lane's point pair in a row
    const int mode = !op.epre2 ? 1 : (op.epre2 == op.epre ? (op.e2f ? 4 : 2) : 3);
    const unsigned o0 = unsigned(rbase) * unsigned(n) + unsigned(c);
    Pair<double> ea[16];
    if (mode == 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const Pair<float> f = ldg_pair(op.e2f, o0 + unsigned(r) * unsigned(n));
            ea[r] = Pair<double>{double(f.x), double(f.y)};
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) ea[r] = ldg_pair(op.epre, o0 + unsigned(r) * unsigned(n));
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        wave_sync();   // the previous reads of the staging rows are done
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int R = 8 * h + r;
            double e0 = ea[R].x, e1 = ea[R].y;
            if (mode == 2) {
                e0 *= e0;
                e1 *= e1;
            } else if (mode == 3) {   // one step per block: the second factor
                const Pair<double> f = ldg_pair(op.epre2, o0 + unsigned(R) * unsigned(n));
                e0 *= f.x;
                e1 *= f.y;
            }
            double2 v;
            v.x = double(xa[R].x) * e0;
            v.y = double(xa[R].y) * e1;
            *reinterpret_cast<double2 *>(&Xw[r * XS + c]) = v;
        }
        wave_sync();
        if ((li >> 3) == h) {
            const double *row = Xw + (li & 7) * XS;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                const int j = lj * NT + k;
                const double x1 = row[j], x2 = row[n - 1 - j];
                sf[k] = x1 + x2;
                af[k] = x1 - x2;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void film_zmma(int rbase, const double (&bs)[16], const double (&ba)[16],
                                          const double (&sf)[16], const double (&af)[16], T *Tm, int lane)
{
    const int li = lane & 15, lj = lane >> 4;
    d4_t U = d4_t{0.0, 0.0, 0.0, 0.0}, W = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        U = __builtin_amdgcn_mfma_f64_16x16x4f64(sf[k], bs[k], U, 0, 0, 0);
        W = __builtin_amdgcn_mfma_f64_16x16x4f64(af[k], ba[k], W, 0, 0, 0);
    }
    // register r of lane l: row l/16 + 4r; output column li of the part
    // (z = k0 + li) and its mirror (z = n-1-k0-li, local column 31 - li)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = rbase + lj + 4 * r;
        Tm[film_at<T>(e, li)] = T(U[r] + W[r]);
        Tm[film_at<T>(e, 31 - li)] = T(U[r] - W[r]);
    }
}

// The same in one piece (loads, staging, products), for the forms without
// the prefetch.
template <typename T>
__device__ __forceinline__ void film_ztile(const FilmOp &op, int rbase, const double (&bs)[16], const double (&ba)[16],
                                           double *Xw, T *Tm, int lane)
{
    constexpr int n = kFilmN, NT = 16, XS = Mfma2<64>::XS;
    const int li = lane & 15, lj = lane >> 4;
    const int c = 2 * lane;   // this lane's point pair in a row
    const T *src = static_cast<const T *>(op.src);
    const int mode = !op.epre2 ? 1 : (op.epre2 == op.epre ? (op.e2f ? 4 : 2) : 3);
    Pair<T> xa[16];
    Pair<double> ea[16];
    const unsigned o0 = unsigned(rbase) * unsigned(n) + unsigned(c);
    if (mode == 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const unsigned a = o0 + unsigned(r) * unsigned(n);
            xa[r] = ldg_pair(src, a);
            const Pair<float> f = ldg_pair(op.e2f, a);
            ea[r] = Pair<double>{double(f.x), double(f.y)};
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const unsigned a = o0 + unsigned(r) * unsigned(n);
            xa[r] = ldg_pair(src, a);
            ea[r] = ldg_pair(op.epre, a);
        }
    }
    double sf[NT], af[NT];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        wave_sync();   // the previous reads of the staging rows are done
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int R = 8 * h + r;
            double e0 = ea[R].x, e1 = ea[R].y;
            if (mode == 2) {
                e0 *= e0;
                e1 *= e1;
            } else if (mode == 3) {   // one step per block: the second factor
                const Pair<double> f = ldg_pair(op.epre2, o0 + unsigned(R) * unsigned(n));
                e0 *= f.x;
                e1 *= f.y;
            }
            double2 v;
            v.x = double(xa[R].x) * e0;
            v.y = double(xa[R].y) * e1;
            *reinterpret_cast<double2 *>(&Xw[r * XS + c]) = v;
        }
        wave_sync();
        if ((li >> 3) == h) {
            const double *row = Xw + (li & 7) * XS;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                const int j = lj * NT + k;
                const double x1 = row[j], x2 = row[n - 1 - j];
                sf[k] = x1 + x2;
                af[k] = x1 - x2;
            }
        }
    }
    d4_t U = d4_t{0.0, 0.0, 0.0, 0.0}, W = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        U = __builtin_amdgcn_mfma_f64_16x16x4f64(sf[k], bs[k], U, 0, 0, 0);
        W = __builtin_amdgcn_mfma_f64_16x16x4f64(af[k], ba[k], W, 0, 0, 0);
    }
    // register r of lane l: row l/16 + 4r; output column li of the part
    // (z = k0 + li) and its mirror (z = n-1-k0-li, local column 31 - li)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int e = rbase + lj + 4 * r;
        Tm[film_at<T>(e, li)] = T(U[r] + W[r]);
        Tm[film_at<T>(e, 31 - li)] = T(U[r] - W[r]);
    }
}

// Transpose cells of one complex line: element (p, k) of the P x B tile
// (lane p, lane-DFT output k) lives in the line's own LDS row
// 16 k + (p ^ (k & 7)); the xor spreads the rows of one access phase over
// the banks both when a column is written (p varies) and when a row is
// read (k varies).
__device__ __forceinline__ int film_cell(int p, int k)
{
    return 16 * k + (p ^ (k & 7));
}

// One exchange through a line's LDS cells: every lane of the line writes
// a[k] to cell wc(k) for k < na, then reads b[j] from cell rc(j) for
// j < nb (the caller's maps are a transpose, so no two lanes write one
// cell).  FP64 rows hold a complex value per cell; FP32 rows hold one
// double, so the real and the imaginary parts go through in two rounds.
template <typename T, int NA, int NB, typename WC, typename RC>
__device__ __forceinline__ void film_xchg(T *Tm, int i, const cpx<double> (&a)[NA], int na, WC wc, cpx<double> (&b)[NB],
                                          int nb, RC rc)
{
    auto ptr = [&](int e) { return reinterpret_cast<double *>(Tm + e * kFilmCols + 2 * film_slot<T>(e, i)); };
    wave_sync();   // earlier reads of these cells are done
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int k = 0; k < NA; ++k)
            if (k < na) *reinterpret_cast<double2 *>(ptr(wc(k))) = double2{a[k].x, a[k].y};
        wave_sync();
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nb) {
                const double2 t = *reinterpret_cast<const double2 *>(ptr(rc(j)));
                b[j] = cpx<double>{t.x, t.y};
            }
    } else {
#pragma unroll
        for (int k = 0; k < NA; ++k)
            if (k < na) ptr(wc(k))[0] = a[k].x;
        wave_sync();
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nb) b[j].x = ptr(rc(j))[0];
        wave_sync();
#pragma unroll
        for (int k = 0; k < NA; ++k)
            if (k < na) ptr(wc(k))[0] = a[k].y;
        wave_sync();
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nb) b[j].y = ptr(rc(j))[0];
    }
}

// The x operator on complex line i of the part (local columns 2i, 2i+1:
// z points z0, z0 + 1) from the part's LDS rows to the output slice.  Lane
// p of the line's 16 holds x[p + 16 m], m < B = L/16: the steps of
// freg_apply<16> (L = 256) or freg2_apply<B, 16>, with the tile transposes
// through the line's own cells.
template <int L, typename T>
__device__ __forceinline__ void film_xops(const double *Kx, T *Tm, const cpx<double> *TW, int i, int p,
                                          cpx<double> (&v)[L / 16])
{
    constexpr int B = L / 16, R = (B + 15) / 16;
    typedef T t2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int m = 0; m < B; ++m) {
        const int e = p + 16 * m;
        const t2 s = *reinterpret_cast<const t2 *>(Tm + e * kFilmCols + 2 * film_slot<T>(e, i));
        v[m] = cpx<double>{double(s.x), double(s.y)};
    }
    // forward: over m, twiddle w^{-p k2} (p k2 < L)
    if constexpr (L == 256) dft16<-1>(v);
    else dft_lane<B, -1>(v);
#pragma unroll
    for (int k = 1; k < B; ++k) {
        const int e = p * k;
        v[k] = cmul(v[k], TW[e + (e >> 4)]);
    }
    // lane q takes the columns k2 = q + 16 r < B of the tile
    const int q = p;
    const int ncol = (q + 16 < B) ? 2 : (q < B ? 1 : 0);
    cpx<double> col[16 * R];
    film_xchg<T>(
        Tm, i, v, B, [&](int k) { return film_cell(p, k); }, col, 16 * ncol,
        [&](int j) { return film_cell(j & 15, q + 16 * (j >> 4)); });
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int c = q + 16 * r;
        if (c < B) {
            dft16<-1>(col, 16 * r);   // over p: spectrum index c + B k1
#pragma unroll
            for (int k1 = 0; k1 < 16; ++k1) {
                const double kk = ldg(Kx, unsigned(c + B * k1)) * (1.0 / L);   // 1/L folded in (exact)
                col[16 * r + k1].x *= kk;
                col[16 * r + k1].y *= kk;
            }
            dft16<+1>(col, 16 * r);
        }
    }
    if constexpr (L == 256) {   // freg_apply twiddles back before the transpose
#pragma unroll
        for (int k = 1; k < 16; ++k) {
            const int e = k * q;
            cpx<double> w = TW[e + (e >> 4)];
            w.y = -w.y;
            col[k] = cmul(col[k], w);
        }
    }
    film_xchg<T>(
        Tm, i, col, 16 * ncol, [&](int j) { return film_cell(j & 15, q + 16 * (j >> 4)); }, v, B,
        [&](int k) { return film_cell(p, k); });
    if constexpr (L != 256) {   // freg2_apply twiddles back after it
#pragma unroll
        for (int k = 1; k < B; ++k) {
            const int e = p * k;
            cpx<double> w = TW[e + (e >> 4)];
            w.y = -w.y;
            v[k] = cmul(v[k], w);
        }
    }
    if constexpr (L == 256) dft16<+1>(v);   // v[m] = x[p + 16 m] after the operator
    else dft_lane<B, +1>(v);
}

// The film step's x operator on complex line i, stored to the output slice.
template <int L, typename T>
__device__ __forceinline__ void film_xline(const FilmOp &op, T *Tm, const cpx<double> *TW, int i, int p, int z0)
{
    constexpr int B = L / 16, n = kFilmN;
    cpx<double> v[B];
    film_xops<L, T>(op.Kx, Tm, TW, i, p, v);
    T *dst = static_cast<T *>(op.dst);
#pragma unroll
    for (int m = 0; m < B; ++m) {
        const unsigned a0 = unsigned(p + 16 * m) * unsigned(n) + unsigned(z0);
        stg_pair(dst, a0, store_round<T>(v[m].x, op.rkey, a0), store_round<T>(v[m].y, op.rkey, a0 + 1));
    }
}

// ------------------------------------- periodic lines through LDS, 32 at a time
//
// The last (periodic) axis of a step for 32 lines whose starts are 32
// consecutive elements (x lines, or y lines of a 3D grid, in groups within
// a row), of length L = 16 B (128, 192, 256, 320).  The workgroup loads
// the 32 columns of its L rows with coalesced 16-byte row-segment loads
// into LDS (the layout of the film step's part rows), runs the film step's
// x transform on its 16 complex lines (4 waves x 4 lines of 16 lanes),
// writes the rounded results back into the same LDS rows, and stores them
// with the same coalesced row segments.  Against the register FFT's
// strided loads (a 32-byte piece of 16 rows per instruction), every memory
// instruction here moves whole 128-byte lines.

template <int L, typename T>
__global__ void __launch_bounds__(64 * kFilmWaves) k_fftx(const LineOp *__restrict__ ops, int gpo)
{
    constexpr int E4 = 16 / int(sizeof(T));        // elements per 16-byte segment
    constexpr int LPR = kFilmCols / E4;            // lanes per row
    constexpr int RPR = 64 * kFilmWaves / LPR;     // rows per round
    static_assert(L % RPR == 0, "rows per round divide the length");
    typedef T tv __attribute__((ext_vector_type(E4)));
    __shared__ __attribute__((aligned(16))) T Tm[L * kFilmCols];
    __shared__ cpx<double> TW[L + L / 16];
    const LineOp &op = ops[blockIdx.x / gpo];
    const int l0 = kFilmCols * (blockIdx.x % gpo);
    if (l0 >= op.nlines) return;   // the whole workgroup: no barrier follows
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double *twd = reinterpret_cast<const double *>(op.tw);
    for (int e = tid; e < L; e += 64 * kFilmWaves)
        TW[e + (e >> 4)] = cpx<double>{ldg(twd, unsigned(2 * e)), ldg(twd, unsigned(2 * e + 1))};
    const unsigned base = unsigned(line_base(op, l0)), pitch = unsigned(op.stride);
    const int rr = tid / LPR, c = E4 * (tid % LPR);   // this thread's row in a round and first column
    const T *src = static_cast<const T *>(op.src);
#pragma unroll
    for (int k = 0; k < L / RPR; ++k) {
        const int e = rr + RPR * k;
        const tv x = *(const __attribute__((address_space(1))) tv *)((const char *)src +
                                                                     (base + unsigned(e) * pitch + unsigned(c)) * sizeof(T));
        *reinterpret_cast<tv *>(Tm + e * kFilmCols + 2 * film_slot<T>(e, c >> 1)) = x;
    }
    __syncthreads();
    const int g = lane & 3, p = lane >> 2;
    const int i = 4 * wave + g;   // complex line: columns 2i, 2i + 1
    cpx<double> v[L / 16];
    film_xops<L, T>(op.K, Tm, TW, i, p, v);
    wave_sync();   // the line's last reads of its cells are done
#pragma unroll
    for (int m = 0; m < L / 16; ++m) {
        const int e = p + 16 * m;
        const unsigned a0 = base + unsigned(e) * pitch + unsigned(2 * i);
        T *cell = Tm + e * kFilmCols + 2 * film_slot<T>(e, i);
        cell[0] = store_round<T>(v[m].x, op.rkey, a0);
        cell[1] = store_round<T>(v[m].y, op.rkey, a0 + 1);
    }
    __syncthreads();
    T *dst = static_cast<T *>(op.dst);
#pragma unroll
    for (int k = 0; k < L / RPR; ++k) {
        const int e = rr + RPR * k;
        const tv x = *reinterpret_cast<const tv *>(Tm + e * kFilmCols + 2 * film_slot<T>(e, c >> 1));
        *(__attribute__((address_space(1))) tv *)((char *)dst + (base + unsigned(e) * pitch + unsigned(c)) * sizeof(T)) = x;
    }
}

// grid: 32 ceil(nops / 8) workgroups; block b runs part (b / 8) % 4 of op
// 8 (b / 32) + b % 8
template <int L, typename T>
__global__ void __launch_bounds__(64 * kFilmWaves, (sizeof(T) == 4 && L <= 256) ? 2 : 1)
    k_film(const FilmOp *__restrict__ ops, int nops, FilmTab tab)
{
    constexpr int NTILE = L / 16, XS = Mfma2<64>::XS;
    __shared__ __attribute__((aligned(16))) T Tm[L * kFilmCols];
    __shared__ __attribute__((aligned(16))) double X[kFilmWaves * 8 * XS];
    __shared__ cpx<double> TW[L + L / 16];
    const int b = blockIdx.x;
    const int o = (b >> 5) * 8 + (b & 7), part = (b >> 3) & 3;
    if (o >= nops) return;   // the whole workgroup: no barrier follows
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const FilmOp &op = ops[o];
    for (int e = tid; e < L; e += 64 * kFilmWaves)
        TW[e + (e >> 4)] = cpx<double>{ldg(tab.tw, unsigned(2 * e)), ldg(tab.tw, unsigned(2 * e + 1))};
    // the part's column of Ms^T / Ma^T for every K step (j = lj 16 + k)
    const int li = lane & 15, lj = lane >> 4;
    double bs[16], ba[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const unsigned a = unsigned((lj * 16 + k) * tab.ldm + 16 * part + li);
        bs[k] = ldg(tab.MsT, a);
        ba[k] = ldg(tab.MaT, a);
    }
    double *Xw = X + wave * 8 * XS;
    // each wave takes NTILE / 4 tiles.  FP32 slices of x length up to 256
    // (two workgroups per CU in 256 registers): the next tile's slice rows
    // load while this tile's products run (past the last tile the last is
    // loaded again); otherwise each tile loads when it starts (the
    // prefetch there would cost a wave per SIMD or spill)
    if constexpr (sizeof(T) == 4 && L <= 256) {
        Pair<T> xa[16];
        film_zload<T>(op, 16 * wave, lane, xa);
        for (int t = wave; t < NTILE; t += kFilmWaves) {
            double sf[16], af[16];
            film_zstage<T>(op, 16 * t, xa, Xw, lane, sf, af);
            film_zload<T>(op, 16 * min(t + kFilmWaves, NTILE - 1), lane, xa);
            film_zmma<T>(16 * t, bs, ba, sf, af, Tm, lane);
        }
    } else {
        for (int t = wave; t < NTILE; t += kFilmWaves) film_ztile<T>(op, 16 * t, bs, ba, Xw, Tm, lane);
    }
    __syncthreads();   // every row of the part's columns is in LDS
    const int g = lane & 3, p = lane >> 2;
    const int i = 4 * wave + g;   // complex line: local columns 2i, 2i + 1
    const int z0 = i < 8 ? 16 * part + 2 * i : kFilmN - kFilmCols - 16 * part + 2 * i;
    film_xline<L, T>(op, Tm, TW, i, p, z0);
}

// ------------------------------------------------------------ noise stream
//
// The reference's deviates (funGaussDev, SCFT.c:1769-1782): POSIX drand48,
// x <- (a x + c) mod 2^48 with a = 0x5DEECE66D, c = 0xB, u = x / 2^48; a
// deviate takes uniform pairs (u1, u2) until w = v1^2 + v2^2 < 1 with
// v = 2u - 1, then returns v1 sqrt(-2 ln w / w) and drops the rest.  One
// iteration draws need = n_active * V deviates (SCFT.c:2103), consuming a
// random number of pairs.  In parallel: candidate pair m of the iteration
// uses draws 2m+1 and 2m+2 from the state at its start; chunk t holds
// pairs [kRngChunk t, kRngChunk (t+1)).  k_rng_count counts accepted pairs
// per chunk (jumping to the chunk's start), k_rng_scan turns the counts
// into the deviate index of each chunk's first accepted pair, k_rng_emit
// writes the deviates in draw order and the state after the pair of the
// last one, and k_rng_commit makes that the next iteration's state.  Each
// instance of a batch has its own stream: grid y runs over the instances
// that draw (RngSlot).  The acceptance test is evaluated with the host's
// IEEE operations (no contraction into FMA), so the accepted pairs, hence
// the state, are the host's bit for bit; the deviate itself goes through
// the device log.

constexpr int kRngChunk = 32;   // candidate pairs per thread
constexpr unsigned long long kLcgA = 0x5DEECE66DULL, kLcgC = 0xBULL, kLcgMask = 0xFFFFFFFFFFFFULL;

struct RngJump {
    unsigned long long a[48], c[48];   // x -> a[b] x + c[b] advances 2^b draws
};

struct RngState {
    unsigned long long x;      // drand48 state at the start of the next iteration
    unsigned long long xn;     // state after this iteration's last deviate
    unsigned long long found;  // accepted pairs among the candidates
    int short_of_pairs;        // the candidates held fewer than `need` accepted pairs
};

// One instance's stream for one iteration.
struct RngSlot {
    RngState *st;
    unsigned *counts;              // [nchunk]
    unsigned long long *offsets;   // [nchunk]
    double *out;                   // need deviates in draw order
    unsigned long long need;
    int nchunk;
};

__device__ __forceinline__ unsigned long long lcg_jump(const RngJump &J, unsigned long long x, unsigned long long k)
{
    for (int b = 0; k; ++b, k >>= 1)
        if (k & 1) x = (J.a[b] * x + J.c[b]) & kLcgMask;
    return x;
}

// One candidate pair from state x (advanced by two draws); returns whether
// it is accepted and sets v1, w.
__device__ __forceinline__ bool polar_pair(unsigned long long &x, double &v1, double &w)
{
#pragma clang fp contract(off)
    x = (kLcgA * x + kLcgC) & kLcgMask;
    const double u1 = double(x) * (1.0 / 281474976710656.0);
    x = (kLcgA * x + kLcgC) & kLcgMask;
    const double u2 = double(x) * (1.0 / 281474976710656.0);
    v1 = 2.0 * u1 - 1.0;
    const double v2 = 2.0 * u2 - 1.0;
    const double a = v1 * v1;
    const double b = v2 * v2;
    w = a + b;
    return w < 1.0;
}

__global__ void __launch_bounds__(kThreads) k_rng_count(const RngSlot *__restrict__ slots,
                                                        const RngJump *__restrict__ J)
{
    const RngSlot &a = slots[blockIdx.y];
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= a.nchunk) return;
    unsigned long long x = lcg_jump(*J, a.st->x, 2ull * kRngChunk * (unsigned long long)t);
    unsigned n = 0;
    for (int m = 0; m < kRngChunk; ++m) {
        double v1, w;
        n += polar_pair(x, v1, w) ? 1u : 0u;
    }
    a.counts[t] = n;
}

// One workgroup per instance: exclusive scan of the chunk counts (each
// thread a contiguous run of chunks, then a workgroup scan of the run totals).
__global__ void __launch_bounds__(1024) k_rng_scan(const RngSlot *__restrict__ slots)
{
    const RngSlot &a = slots[blockIdx.x];
    RngState *st = a.st;
    const unsigned *counts = a.counts;
    unsigned long long *offsets = a.offsets;
    const int nchunk = a.nchunk;
    const unsigned long long need = a.need;
    __shared__ unsigned long long part[1024];
    const int t = threadIdx.x;
    const int per = (nchunk + 1023) / 1024;
    const int b0 = min(t * per, nchunk), b1 = min(b0 + per, nchunk);
    unsigned long long sum = 0;
    for (int i = b0; i < b1; ++i) sum += counts[i];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive Hillis-Steele scan
        const unsigned long long v = t >= d ? part[t - d] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned long long run = part[t] - sum;
    for (int i = b0; i < b1; ++i) {
        offsets[i] = run;
        run += counts[i];
    }
    if (t == 1023) {
        st->found = part[1023];
        st->short_of_pairs = part[1023] < need ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kThreads) k_rng_emit(const RngSlot *__restrict__ slots,
                                                       const RngJump *__restrict__ J)
{
    const RngSlot &a = slots[blockIdx.y];
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= a.nchunk) return;
    const unsigned long long need = a.need;
    unsigned long long j = a.offsets[t];
    if (j >= need || a.counts[t] == 0) return;
    unsigned long long x = lcg_jump(*J, a.st->x, 2ull * kRngChunk * (unsigned long long)t);
    for (int m = 0; m < kRngChunk && j < need; ++m) {
        double v1, w;
        if (!polar_pair(x, v1, w)) continue;
        a.out[j] = v1 * sqrt((-2.0 * log(w)) / w);
        if (j == need - 1) a.st->xn = x;
        ++j;
    }
}

// grid: one workgroup per instance
__global__ void k_rng_commit(const RngSlot *__restrict__ slots)
{
    RngState *st = slots[blockIdx.x].st;
    if (threadIdx.x == 0 && !st->short_of_pairs) st->x = st->xn;
}

// ------------------------------------------------------------------ prep

struct PrepArgs {
    const double *w;     // 13*V
    double *E;           // 13*V
    float *E2f;          // 13*V: E^2 rounded once to FP32 (FP32-split z operator), or null
    void *q0_dbk1;       // dbk1 q slice 0 (T), or null
    int fields[13];      // fields whose Boltzmann factor is needed
    int nfields;
    double ds;
    int V;
    // dbk1 seed start (SCFT.c:2346-2367): mode 0 none, 1 seed box, 2 ones
    int seed_mode;
    int nx, ny, nz, delta;
};

// E^2 rounded to FP32 without bias.  Round-to-nearest gives every voxel
// of a region where w is uniform the same rounding error, which the z
// operator then applies at every step of a block: a systematic relative
// error of up to NS x 2^-25 in q there.  Here the value goes to one of its
// two FP32 neighbours with probability proportional to its closeness, the
// choice made by a hash of the array offset: the rounding errors have zero
// mean and no correlation between voxels, so averages over voxels (Q, the
// free energy) do not accumulate them.
__device__ __forceinline__ float round_unbiased_f32(double x, unsigned key)
{
    const float f = float(x);
    const float lo = (double(f) <= x) ? f : nextafterf(f, 0.0f);
    const float hi = (double(f) <= x) ? nextafterf(f, 3.0e38f) : f;
    if (double(lo) == x) return lo;
    unsigned h = key * 0x9E3779B1u;              // integer hash (multiply-xorshift)
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    const double u = (double(h) + 0.5) * (1.0 / 4294967296.0);
    return (u * (double(hi) - double(lo)) < x - double(lo)) ? hi : lo;
}

// grid: (blocks, instances)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_prep(const PrepArgs *__restrict__ args)
{
    const PrepArgs &a = args[blockIdx.y];
    T *q0 = static_cast<T *>(a.q0_dbk1);
    const int stride = gridDim.x * blockDim.x;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < a.V; idx += stride) {
        for (int i = 0; i < a.nfields; ++i) {
            const size_t o = (size_t)a.fields[i] * a.V + idx;
            const double e = exp(a.ds * (-a.w[o] / 2.0));
            a.E[o] = e;
            if (a.E2f) a.E2f[o] = round_unbiased_f32(e * e, unsigned(o));
        }
        if (a.seed_mode == 2) {
            q0[idx] = T(1);
        } else if (a.seed_mode == 1) {
            const int k = idx % a.nz;
            const int j = (idx / a.nz) % a.ny;
            const int i = idx / (a.nz * a.ny);
            const int D = a.delta;
            const bool in = i > (a.nx / 2 - D) && i < (a.nx / 2 + D) && j > (a.ny / 2 - D) &&
                            j < (a.ny / 2 + D) && k > (a.nz / 4 - D) && k < (a.nz / 4 + D);
            q0[idx] = in ? T(1) : T(0);
        }
    }
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of NV values; result valid in thread 0.
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

// Per-instance scalar block (device).
struct Scalars {
    double Q[17];
    double H;        // F before the ln Q terms
    double F;
    double rms;
    double pmean;
    double dw_rms;
    double lnq_frac;
};

struct QSlot {
    const void *end;   // q(NS_p - 1) slice (T), stored without its factor
    const double *e;   // Boltzmann factor owed by that slice, or null
    double *partial;   // [nblocks]
    int V;
};

// grid: (nblocks, slots of all instances)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_qsum(const QSlot *__restrict__ slots)
{
    const QSlot &s = slots[blockIdx.y];
    const T *q = static_cast<const T *>(s.end);
    double v[1] = {0.0};
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < s.V; idx += gridDim.x * kThreads)
        v[0] += s.e ? double(q[idx]) * s.e[idx] : double(q[idx]);
    block_sum<1>(v);
    if (threadIdx.x == 0) s.partial[blockIdx.x] = v[0];
}

// ------------------------------------------------ densities + F + RMSDev

// Slices are stored without the factor of the step that made them: slice s
// of a chain owes E_a for 1 <= s <= switch_at, E_b above, nothing at s = 0.
struct DenTerm {
    const void *q;
    const void *qd;
    const double *eqa, *eqb, *eda, *edb;   // factors of the q and q-dagger chains
    int qsw, dsw;                          // their switch slices
    int ns, start, stop;
    int field;   // 0..12
    int sp;      // species (index into Scalars::Q)
    double frac; // volume factor (vol (1 - coating) or vol_s); <= 0: not applied
};

struct DenArgs {
    DenTerm term[kMaxTerms];
    int nterms;
    int V;
    int mask;
    int nphi;            // fields carrying density
    int phif[13];
    double chi[13][13];
    const Scalars *scal;
    const double *w;     // 13*V
    const double *p;     // V
    double *tot;         // 13*V
    double *phitot;      // V
    double *partial;     // [2][nblocks]: H, (phi_tot-1)^2
};

template <typename T>
__device__ __forceinline__ double simpson(const DenTerm &t, int V, int idx)
{
    const size_t sV = (size_t)V;
    const int top = t.ns - 1;
    const T *q = static_cast<const T *>(t.q);
    const T *qd = static_cast<const T *>(t.qd);
    const unsigned ui = unsigned(idx);
    const double qa = ldg(t.eqa, ui), qb = ldg(t.eqb, ui), da = ldg(t.eda, ui), db = ldg(t.edb, ui);
    auto fac = [](int s, int sw, double a, double b) { return s == 0 ? 1.0 : (s <= sw ? a : b); };
    // slice base pointers are uniform (scalar arithmetic), the point offset is
    // 32-bit; every slice is read once per iteration (streaming loads)
    auto qc = [&](int s) -> double {
        const int u = top - s;
        return (double(ldg_nt(q + (size_t)s * sV, ui)) * fac(s, t.qsw, qa, qb)) *
               (double(ldg_nt(qd + (size_t)u * sV, ui)) * fac(u, t.dsw, da, db));
    };
    if constexpr (sizeof(T) == 8) {
        // the reference's order: ends, then 4*odd, then 2*even (SCFT.c:2283-2297)
        double phi = qc(t.start) + qc(t.stop);
        for (int s = t.start + 1; s <= t.stop - 1; s += 2) phi = phi + 4.0 * qc(s);
        for (int s = t.start + 2; s <= t.stop - 2; s += 2) phi = phi + 2.0 * qc(s);
        return phi;
    } else {
        // FP32 slices: products and the contour sum in FP64 with Kahan compensation
        double sum = qc(t.start) + qc(t.stop), comp = 0.0;
        auto add = [&](double y) {
            y -= comp;
            const double tt = sum + y;
            comp = (tt - sum) - y;
            sum = tt;
        };
        if (t.q == t.qd && t.start == 0 && t.stop == top && !(top & 1)) {
            // one chain over the whole contour (homopolymer): qc(s) = qc(top - s)
            // and the weights are symmetric, so each slice is read once
#pragma unroll 8
            for (int s = 1; s < top / 2; ++s) add(((s & 1) ? 8.0 : 4.0) * qc(s));
            if (top > 0) add((((top / 2) & 1) ? 4.0 : 2.0) * qc(top / 2));
        } else {
#pragma unroll 8
            for (int s = t.start + 1; s <= t.stop - 1; ++s) add((((s - t.start) & 1) ? 4.0 : 2.0) * qc(s));
        }
        return sum;
    }
}

// grid: (nblocks, instances)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_density(const DenArgs *__restrict__ args)
{
    const DenArgs &a = args[blockIdx.y];
    const int V = a.V;
    double acc[2] = {0.0, 0.0};
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads) {
        double phi[13];
        bool touched[13];
        for (int f = 0; f < 13; ++f) { phi[f] = 0.0; touched[f] = false; }
        for (int i = 0; i < a.nterms; ++i) {
            const DenTerm &t = a.term[i];
            const double dseff = 1.0 / (t.ns - 1);
            double v = simpson<T>(t, V, idx) / 3.0;
            v = v * (dseff / a.scal->Q[t.sp]);
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
    double *partial;     // [2][nblocks]: sum p_new, sum dw^2
};

// funUpdate for the voxels of one thread (SCFT.c:2047-2113), with NA the
// number of active fields when it is known at compile time (the loops then
// unroll and the per-voxel arrays stay in registers) or 0 (a.nact)
template <int NA>
__device__ __forceinline__ void update_voxels(const UpdArgs &a, double (&acc)[2])
{
    const int V = a.V;
    const int n = NA > 0 ? NA : a.nact;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < V; idx += gridDim.x * kThreads) {
        double w0[NA > 0 ? NA : 13], r[NA > 0 ? NA : 13], dW[NA > 0 ? NA : 13];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            const int f = a.act[i];
            w0[i] = a.w[(size_t)f * V + idx];
            r[i] = a.tot[(size_t)f * V + idx] - a.phibar[f];
        }
        const double pv = a.p[idx];
#pragma unroll
        for (int i = 0; i < n; ++i) {
            double d = -w0[i] + pv;
#pragma unroll
            for (int j = 0; j < n; ++j)
                if (j != i) d += a.chi[a.act[i]][a.act[j]] * r[j];
            dW[i] = d;
        }
        double corr = 0.0;
        if (n > 1)
#pragma unroll
            for (int i = 0; i < n; ++i)
#pragma unroll
                for (int j = i + 1; j < n; ++j) {
                    const double c = a.chi[a.act[i]][a.act[j]];
#pragma unroll
                    for (int h = 0; h < n; ++h)
                        if (h != i && h != j) corr += c * r[h];
                }
        size_t draw = 0;
        if (a.noise) {
            const int k = idx % a.nz, j = (idx / a.nz) % a.ny, i = idx / (a.nz * a.ny);
            draw = ((size_t)(k * a.ny + j) * a.nx + i) * n;
        }
        double psum = 0.0;
#pragma unroll
        for (int i = 0; i < n; ++i) {
            double wn = w0[i];
#pragma unroll
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
}

// funUpdate per voxel (SCFT.c:2047-2113); grid: (nblocks, instances)
__global__ void __launch_bounds__(kThreads) k_update(const UpdArgs *__restrict__ args)
{
    const UpdArgs &a = args[blockIdx.y];
    double acc[2] = {0.0, 0.0};
    switch (a.nact) {   // uniform per workgroup
    case 2: update_voxels<2>(a, acc); break;
    case 3: update_voxels<3>(a, acc); break;
    case 4: update_voxels<4>(a, acc); break;
    case 5: update_voxels<5>(a, acc); break;
    case 6: update_voxels<6>(a, acc); break;
    default: update_voxels<0>(a, acc); break;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        a.partial[blockIdx.x] = acc[0];
        a.partial[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ------------------------------------------------------------- finalize

struct FinArgs {
    int V, nblocks, nq, nact;
    int qsp[17];          // species of each Q slot
    double lnq_coef[17];  // vol/alpha * c per Q slot (SCFT.c:1529-1555)
    const double *qpart;  // [nq][nblocks]
    const double *dpart;  // [2][nblocks]
    const double *upart;  // [2][nblocks]
    double *p;            // V
    Scalars *out;
};

// Q from its partials; grid: (instances), block (kThreads).  Up to 256
// partials: thread s < nq sums slot s in order.  More (grids of more than
// 64 K voxels): wave w sums slots w, w + 4, ... across its lanes, so that
// no thread adds 1024 partials one after another.
__global__ void __launch_bounds__(kThreads) k_qfinal(const FinArgs *__restrict__ args)
{
    const FinArgs &a = args[blockIdx.x];
    if (a.nblocks <= 256) {
        for (int s = threadIdx.x; s < a.nq; s += kThreads) {
            double v = 0.0;
            for (int b = 0; b < a.nblocks; ++b) v += a.qpart[s * a.nblocks + b];
            a.out->Q[a.qsp[s]] = v / a.V;
        }
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int s = wave; s < a.nq; s += kThreads / 64) {
        double v = 0.0;
        for (int b = lane; b < a.nblocks; b += 64) v += a.qpart[s * a.nblocks + b];
        v = wave_sum(v);
        if (lane == 0) a.out->Q[a.qsp[s]] = v / a.V;
    }
}

// F, RMSDev, mean p, ||dw||; grid: (instances)
__global__ void __launch_bounds__(kThreads) k_final(const FinArgs *__restrict__ args)
{
    const FinArgs &a = args[blockIdx.x];
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

// p -= mean(p) (SCFT.c:2119-2123); grid: (blocks, instances)
__global__ void __launch_bounds__(kThreads) k_pshift(const FinArgs *__restrict__ args)
{
    const FinArgs &a = args[blockIdx.y];
    const double m = a.out->pmean;
    for (int idx = blockIdx.x * kThreads + threadIdx.x; idx < a.V; idx += gridDim.x * kThreads) a.p[idx] -= m;
}

// ------------------------------------------------------------- utility

template <typename T>
__global__ void __launch_bounds__(kThreads) k_convert(const double *src, T *dst, int n)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) dst[i] = T(src[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_widen(const T *src, const double *e, double *dst, int n)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
        dst[i] = e ? double(src[i]) * e[i] : double(src[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill(T *dst, T v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) dst[i] = v;
}

}  // namespace scft
