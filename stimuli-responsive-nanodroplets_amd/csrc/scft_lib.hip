// scft_lib.hip — libscft: the SCFT.c iteration loop on one MI355X (gfx950).
//
// Host side of the C ABI declared in include/scft.h.  A context is a batch
// of SCFT instances ("slots") that share one chain structure: the same
// species set, NS, MASK, NY, NZ, LY, LZ and DR_TARGET; chi, volumes of the
// same species set, G1 and LX (hence NX) may differ per slot.  Everything
// an iteration touches is device-resident: fields, Boltzmann factors,
// propagators (stored [s][idx], one contiguous slice per contour point),
// densities and per-slot scalars.  One iteration of every active slot is a
// fixed launch sequence on the context's stream:
//   k_prep       E_f = exp(-DS w_f / 2) for every used field (+ dbk1 seed)
//   per contour step t, per axis (z, y, x): one k_zdense / k_lineop launch
//                whose ops are every chain of every active slot that still
//                has a step t (SCFT.c:1092-1456 block structure)
//   k_qsum, k_qfinal   Q_p = mean q_p(NS_p - 1)                      (FP64)
//   k_density    Simpson contour sums, volume scaling, totals, free-energy
//                integrand and (phi_tot - 1)^2
//   k_update     funUpdate per voxel                                   (FP64)
//   k_final, k_pshift  F, RMSDev, mean p, ||dw||; p -= mean(p)
//   one D2H copy of every slot's scalars
// The sequence is captured once as a hipGraph and replayed while the active
// set does not change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/scft.h"
#include "scft_kernels.h"

using namespace scft;

namespace {

thread_local std::string g_err;

void set_err(const std::string &s) { g_err = s; }

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            set_err(std::string(#x) + ": " + hipGetErrorString(e_));                      \
            return SCFT_EHIP;                                                             \
        }                                                                                 \
    } while (0)

#define RETCHK(x)                      \
    do {                               \
        int r_ = (x);                  \
        if (r_ != SCFT_OK) return r_;  \
    } while (0)

const int kDbkField[4][2] = {{0, 1}, {3, 4}, {5, 6}, {7, 8}};

struct Radices {
    int npass = 0;
    int r[kMaxPass];
};

bool plan_radices(int L, Radices &pl)
{
    int n = L;
    pl.npass = 0;
    auto push = [&](int R) {
        if (pl.npass >= kMaxPass) return false;
        pl.r[pl.npass++] = R;
        n /= R;
        return true;
    };
    while (n % 4 == 0) if (!push(4)) return false;
    while (n % 2 == 0) if (!push(2)) return false;
    while (n % 3 == 0) if (!push(3)) return false;
    while (n % 5 == 0) if (!push(5)) return false;
    for (int p = 7; p * p <= n; p += 2)
        while (n % p == 0) if (!push(p)) return false;
    if (n > 1) return push(n);
    return true;
}

// K of one axis on a transform grid of length L (SCFT.c:983-1023 factorised):
// exp(-alpha k_m^2), k_m = 2 pi (L/2 - |m - L/2|) / L.
double axis_k(int m, int L, double alpha)
{
    const double km = 2.0 * M_PI * (L / 2.0 - std::fabs(m - L / 2.0)) / L;
    return std::exp(-alpha * std::pow(km, 2));
}

// The z operator as a dense matrix, M[a][b] = response at a to a unit
// impulse at b, from the eigen-decomposition of the reference's transform:
//   periodic (length n):  (1/n) sum_m K_m cos(2 pi m (a - b) / n)
//   Dirichlet (N = 2(n+1)): (2/(n+1)) sum_{m=1..n} K_m sin(pi m (a+1)/(n+1)) sin(pi m (b+1)/(n+1))
// summed in long double.
std::vector<double> dense_operator(int n, int dirichlet, double alpha)
{
    std::vector<double> M((size_t)n * n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            long double s = 0.0L;
            if (dirichlet) {
                const int N = 2 * (n + 1);
                for (int m = 1; m <= n; ++m)
                    s += (long double)axis_k(m, N, alpha) * sinl(pi * m * (a + 1) / (n + 1)) * sinl(pi * m * (b + 1) / (n + 1));
                s *= 2.0L / (n + 1);
            } else {
                for (int m = 0; m < n; ++m) {
                    const long e = ((long)m * (long)(a - b + n)) % n;
                    s += (long double)axis_k(m, n, alpha) * cosl(2.0L * pi * e / n);
                }
                s /= n;
            }
            M[(size_t)a * n + b] = (double)s;
        }
    return M;
}

template <typename X>
int dev_alloc(X **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), n * sizeof(X));
    if (e != hipSuccess) {
        set_err(std::string("hipMalloc: ") + hipGetErrorString(e));
        return SCFT_ENOMEM;
    }
    return SCFT_OK;
}

int grid_for(size_t n, int cap = 2048)
{
    size_t g = (n + kThreads - 1) / kThreads;
    return int(std::max<size_t>(1, std::min<size_t>(g, cap)));
}

// Device-side per-size tables, shared by the slots of a context.
struct FftTable {
    int L = 0;
    Radices rad;
    double *K = nullptr;
    cpx<double> *tw = nullptr;
};
struct DenseTable {
    int n = 0, hs = 0, ha = 0, ldm = 0;
    double *MsT = nullptr, *MaT = nullptr;
};

enum ZKind { Z_FFT = 0, Z_DENSE = 1 };

// Smallest grid (points per instance) whose FP32-split z operator reads
// E^2 as FP32.
constexpr int kE2fMinV = 16384;

// Process-wide autotuner results: (n, dirichlet, alpha bits, storage,
// slots timed, NX, candidate set) -> candidate (autotune_z).
std::mutex g_tune_mu;
std::map<std::tuple<int, int, long long, int, int, int, int, int, int>, int> g_tune;

// The autotuner times as many slots as put this many voxels in one launch
// (at most the batch): one small slot alone times launch latency only.
constexpr long long kTuneVoxels = 1LL << 22;

}  // namespace

// ---------------------------------------------------------------- context

// A propagator chain: step t (slice t -> t+1) uses field fa while
// t < switch_at, fb afterwards.  Slices are stored without the Boltzmann
// factor of the step that produced them (q(s) = E(s) * stored(s), E(0) = 1);
// the first axis of the next step applies it together with its own.
struct Chain {
    void *base;       // q or q-dagger array of the species (T)
    void *tmp;        // V values between axis operators (T, or FP64 when wide)
    int nsteps, fa, fb, switch_at;
    int step_field(int t) const { return t < switch_at ? fa : fb; }
    // field of the factor owed by stored slice s, or -1 (slice 0)
    int slice_field(int s) const { return s == 0 ? -1 : step_field(s - 1); }
};

struct Slot {
    scft_params p{};
    scft_derived d{};
    int NX = 0, NY = 0, NZ = 0, V = 0;
    bool ready = false, active = false;
    unsigned long iter = 0;
    double *w = nullptr, *pf = nullptr, *tot = nullptr, *phitot = nullptr, *mask = nullptr, *E = nullptr;
    float *E2f = nullptr;   // E^2 in FP32 for the z operator inside a block (FP32-split), or null
    void *q[SCFT_NSPECIES] = {};
    void *qd[SCFT_NSPECIES] = {};
    double *qpart = nullptr, *dpart = nullptr, *upart = nullptr;
    double *noise = nullptr;
    // device noise stream (scft_slot_rng_seed): candidate chunks and their scan
    bool rng_on = false;
    unsigned *rng_counts = nullptr;
    unsigned long long *rng_offsets = nullptr;
    int rng_nchunk = 0, rng_cap = 0;
    unsigned long long rng_need = 0;     // deviates per iteration (n_active * V)
    std::vector<Chain> chains;
    const FftTable *fx = nullptr, *fy = nullptr, *fz = nullptr;
    const DenseTable *dz = nullptr;
    PrepArgs prep{};
    DenArgs den{};
    UpdArgs upd{};
    FinArgs fin{};
    std::vector<QSlot> qslots;
};

enum LaunchKind {
    LK_LINEOP = 0, LK_DENSE = 1, LK_FFTWAVE = 2, LK_DENSE_MFMA = 3, LK_DENSE_MFMA2 = 4, LK_FFTREG = 5, LK_FFTREG2 = 6,
    LK_PLANE = 7, LK_FILM = 8, LK_FFTX = 9, LK_FFTMIX = 10
};
constexpr int kLaunchKinds = 11;

struct Launch {
    int kind;        // LaunchKind
    int pos;         // 0 first, 1 middle, 2 last, 3 only
    int first, count, P, gx;
    int L = 0;       // transform length (FFT kinds)
    int hsp = 0;     // padded dense half width (MFMA kind)
    size_t lds;
    FilmTab ftab{};  // tables of a fused film step (LK_FILM)
    int step = 0;    // contour step
    bool pf = false; // LK_DENSE_MFMA2: the kernel form with the next tile's slice rows in flight
    MixArgs mix{};   // LK_FFTMIX: the length groups
    ZTiles zt{};     // LK_DENSE_MFMA2: tiles of the ops grouped by row count
    int ntiles = 0;  // LK_DENSE_MFMA2: tiles of the launch
};

bool fftwave_supported(int L)
{
    switch (L) {
    case 16: case 32: case 48: case 64: case 80: case 96: case 128: case 160:
    case 192: case 256: case 320: case 384: case 512: return true;
    default: return false;
    }
}

struct scft_ctx {
    int precision = SCFT_FP64;
    size_t tsize = 8;                // sizeof(T)
    scft_params proto{};
    scft_derived dproto{};
    int nslots = 0, max_nx = 0, max_V = 0;
    int NY = 0, NZ = 0;
    hipStream_t stream = nullptr;
    // launches of one step and axis position (one per x length) run side by side
    // on these streams, forked from and joined back into the context's stream
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> evs;
    bool par_axes = true;            // SCFT_PARX=0: one after another
    bool xmix = true;                // one launch for the x lengths 192/256/320 of a step (SCFT_XMIX=0: off)
    bool zpf = true;                 // z operator with the next tile's slice rows in flight (SCFT_ZPF=0: off)
    std::vector<Slot> slots;
    std::vector<void *> allocs;
    std::vector<std::unique_ptr<FftTable>> ffts;
    std::vector<std::unique_ptr<DenseTable>> denses;
    int zkind = Z_FFT;
    bool fft_wave = true;            // wave-per-line FFT for supported periodic lengths
    bool fft_reg = true;             // register FFTs for periodic lengths 64, 96, 128, 192, 256, 320
    bool fft_reg_force = false;      // register FFTs also for launches of few lines
    int fftx_mode = 2;               // periodic lines through LDS in 32-line groups (k_fftx):
                                     // 0 never, 1 every eligible length, 2 autotuned per size class
    bool xcoal_cls[3] = {};          // k_fftx chosen for active sets of <= 16, <= 64, more slots
    bool dense_mfma = true;          // MFMA dense z operator when hs <= 64
    bool dense_mfma1 = false;        // force the one-wave-per-SIMD MFMA kernel for even NZ too
    bool tmp_wide = false;           // FP64 scratch between axis operators in FP32 mode
    bool e2f = false;                // FP32 E^2 for the MFMA z operator inside a block (FP32 mode)
    bool sround = false;             // FP32 slices and scratch rounded without bias (store_round)
    int ncu = 256;                   // compute units of the device (persistent launches)
    Scalars *scal = nullptr;         // [nslots] device
    Scalars *scal_host = nullptr;    // [nslots] pinned
    // schedule of the active set
    bool dirty = true;
    std::vector<int> act;            // active slot ids
    std::vector<Launch> launches;
    LineOp *lineops = nullptr;
    size_t lineops_cap = 0;
    DenseOp *denseops = nullptr;
    size_t denseops_cap = 0;
    PlaneOp *planeops = nullptr;
    size_t planeops_cap = 0;
    bool plane_on = true;            // fused step on periodic 64 x 64 planes (SCFT_PLANE=0: off)
    FilmOp *filmops = nullptr;
    size_t filmops_cap = 0;
    int film_mode = -1;              // fused z + x step on 2D films: 0 off, 1 on, -1 autotuned (SCFT_FILM)
    bool film_cls[3] = {};           // film step chosen for active sets of <= 16, <= 64, more slots
    int nstep0 = 0;                  // launches of contour step 0
    std::vector<int> step_first;     // index of each contour step's first launch
    bool inflight = false;           // a batched iteration launched by scft_batch_launch, not yet completed
    PrepArgs *d_prep = nullptr;
    DenArgs *d_den = nullptr;
    UpdArgs *d_upd = nullptr;
    FinArgs *d_fin = nullptr;
    QSlot *d_qslots = nullptr;
    size_t d_qslots_cap = 0;
    int nqslots = 0;
    int nred = 1;
    // graph of one iteration of the active set
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // device replay of the noise stream, one drand48 state per slot
    RngState *rng = nullptr;             // [nslots] device
    RngState *rng_host = nullptr;        // [nslots] pinned copy, refreshed with the scalars
    RngJump *rng_jump = nullptr;         // device
    RngSlot *d_rng = nullptr;            // this schedule's drawing slots
    size_t d_rng_cap = 0;
    int nrng = 0, rng_maxchunk = 0;
};

namespace {

template <typename X>
int ctx_alloc(scft_ctx *c, X **p, size_t n)
{
    RETCHK(dev_alloc(p, n));
    c->allocs.push_back(*p);
    return SCFT_OK;
}

void fill_d(scft_ctx *c, double *p, double v, size_t n)
{
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, p, v, n);
}

void fill_t(scft_ctx *c, void *p, double v, size_t n)
{
    if (c->tsize == 8)
        hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, (double *)p, v, n);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(kThreads), 0, c->stream, (float *)p, (float)v, n);
}

// Slots must share everything that shapes the launch schedule.
bool same_structure(const scft_params &a, const scft_params &b)
{
    if (a.ns != b.ns || a.mask != b.mask || a.ly != b.ly || a.lz != b.lz || a.dr_target != b.dr_target) return false;
    if ((a.lx == 1.0) != (b.lx == 1.0)) return false;   // every slot has an x axis, or none has
    for (int i = 0; i < 4; ++i) {
        if ((a.vol_dbk[i] > 0) != (b.vol_dbk[i] > 0)) return false;
        if (a.vol_dbk[i] > 0 && (a.len_dbk[i] != b.len_dbk[i] || a.f_dbk[i] != b.f_dbk[i])) return false;
    }
    for (int i = 0; i < 10; ++i) {
        if ((a.vol_homo[i] > 0) != (b.vol_homo[i] > 0)) return false;
        if (a.vol_homo[i] > 0 && a.len_homo[i] != b.len_homo[i]) return false;
    }
    for (int i = 0; i < 3; ++i) {
        if ((a.vol_s[i] > 0) != (b.vol_s[i] > 0)) return false;
        if (a.vol_s[i] > 0 && a.len_s[i] != b.len_s[i]) return false;
    }
    return true;
}

int get_fft(scft_ctx *c, int L, double alpha, const FftTable **out)
{
    auto t = std::make_unique<FftTable>();
    t->L = L;
    if (!plan_radices(L, t->rad)) {
        set_err("transform length " + std::to_string(L) + " has too many factors");
        return SCFT_EUNSUPPORTED;
    }
    if (size_t(2) * L * sizeof(cpx<double>) > 160 * 1024) {
        set_err("transform length " + std::to_string(L) + " exceeds the LDS tile");
        return SCFT_EUNSUPPORTED;
    }
    std::vector<double> K(L);
    std::vector<cpx<double>> tw(L);
    for (int m = 0; m < L; ++m) {
        K[m] = axis_k(m, L, alpha);
        const double a = 2.0 * M_PI * double(m) / double(L);
        tw[m] = {std::cos(a), -std::sin(a)};
    }
    RETCHK(ctx_alloc(c, &t->K, L));
    RETCHK(ctx_alloc(c, &t->tw, L));
    HIPCHK(hipMemcpy(t->K, K.data(), L * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->tw, tw.data(), L * sizeof(cpx<double>), hipMemcpyHostToDevice));
    *out = t.get();
    c->ffts.push_back(std::move(t));
    return SCFT_OK;
}

int get_dense(scft_ctx *c, int n, int dirichlet, double alpha, const DenseTable **out)
{
    auto t = std::make_unique<DenseTable>();
    t->n = n;
    t->ha = n / 2;
    t->hs = n - t->ha;
    t->ldm = ((t->hs + 3) / 4) * 4;
    std::vector<double> M = dense_operator(n, dirichlet, alpha);
    std::vector<double> MsT((size_t)t->hs * t->ldm, 0.0), MaT((size_t)std::max(1, t->ha) * t->ldm, 0.0);
    // Ms[k][j] = (M[k][j] + M[k][n-1-j]) / 2 (j < ha), Ms[k][mid] = M[k][mid];
    // Ma[k][j] = (M[k][j] - M[k][n-1-j]) / 2; stored transposed.
    for (int k = 0; k < t->hs; ++k) {
        for (int j = 0; j < t->ha; ++j) {
            MsT[(size_t)j * t->ldm + k] = 0.5 * (M[(size_t)k * n + j] + M[(size_t)k * n + n - 1 - j]);
            MaT[(size_t)j * t->ldm + k] = 0.5 * (M[(size_t)k * n + j] - M[(size_t)k * n + n - 1 - j]);
        }
        if (t->hs > t->ha) MsT[(size_t)t->ha * t->ldm + k] = M[(size_t)k * n + t->ha];
    }
    RETCHK(ctx_alloc(c, &t->MsT, MsT.size()));
    RETCHK(ctx_alloc(c, &t->MaT, MaT.size()));
    HIPCHK(hipMemcpy(t->MsT, MsT.data(), MsT.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t->MaT, MaT.data(), MaT.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = t.get();
    c->denses.push_back(std::move(t));
    return SCFT_OK;
}

int lines_P(int L)
{
    const int budget = 64 * 1024;
    int P = budget / int(2 * L * sizeof(cpx<double>));
    return std::max(1, std::min(P, 16));
}

// ------------------------------------------------------------- slot setup

int slot_alloc(scft_ctx *c, Slot &s)
{
    const size_t V = c->max_V;
    const scft_derived &d = c->dproto;
    RETCHK(ctx_alloc(c, &s.w, 13 * V));
    RETCHK(ctx_alloc(c, &s.pf, V));
    RETCHK(ctx_alloc(c, &s.tot, 13 * V));
    RETCHK(ctx_alloc(c, &s.phitot, V));
    RETCHK(ctx_alloc(c, &s.mask, 3 * V));
    RETCHK(ctx_alloc(c, &s.E, 13 * V));
    if (c->e2f) RETCHK(ctx_alloc(c, &s.E2f, 13 * V));
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        const int ns = d.ns_sp[sp];
        if (ns <= 0) continue;
        char *q = nullptr;
        RETCHK(ctx_alloc(c, &q, (size_t)ns * V * c->tsize));
        s.q[sp] = q;
        if (sp >= 4 && sp < 14) {
            s.qd[sp] = q;   // q-dagger = q for homopolymers (SCFT.c:1102)
        } else {
            char *qd = nullptr;
            RETCHK(ctx_alloc(c, &qd, (size_t)ns * V * c->tsize));
            s.qd[sp] = qd;
        }
    }
    RETCHK(ctx_alloc(c, &s.qpart, (size_t)SCFT_NSPECIES * kRedBlocks));
    RETCHK(ctx_alloc(c, &s.dpart, 2 * (size_t)kRedBlocks));
    RETCHK(ctx_alloc(c, &s.upart, 2 * (size_t)kRedBlocks));
    // chains in the reference's species order (SCFT.c:1092-1456)
    s.chains.clear();
    auto add = [&](void *base, int nsteps, int fa, int fb, int sw) {
        s.chains.push_back(Chain{base, nullptr, nsteps, fa, fb, sw});
    };
    for (int h = 0; h < 10; ++h)
        if (d.ns_sp[4 + h] > 0) add(s.q[4 + h], d.ns_sp[4 + h] - 1, h, h, 1 << 30);
    for (int b = 0; b < 4; ++b) {
        if (d.ns_sp[b] <= 0) continue;
        const int fa = kDbkField[b][0], fb = kDbkField[b][1];
        add(s.q[b], d.ns_sp[b] - 1, fa, fb, d.ns_x[b] - 1);   // A then B (SCFT.c:1250-1256)
        add(s.qd[b], d.ns_sp[b] - 1, fb, fa, d.ns_y[b]);      // B then A (SCFT.c:1263-1269)
    }
    for (int b = 0; b < 3; ++b) {
        const int sp = 14 + b;
        if (d.ns_sp[sp] <= 0) continue;
        add(s.q[sp], d.ns_sp[sp] - 1, 10 + b, 10 + b, 1 << 30);   // from the mask (SCFT.c:2441)
        add(s.qd[sp], d.ns_sp[sp] - 1, 10 + b, 10 + b, 1 << 30);
    }
    for (auto &ch : s.chains) {
        double *t = nullptr;
        RETCHK(ctx_alloc(c, &t, V));
        ch.tmp = t;
    }
    return SCFT_OK;
}

const Chain *find_chain(const Slot &s, const void *base)
{
    for (auto &ch : s.chains)
        if (ch.base == base) return &ch;
    return nullptr;
}

// Boltzmann factor owed by stored slice idx of a chain (null for slice 0).
const double *slice_factor(const Slot &s, const Chain &ch, int idx)
{
    const int f = ch.slice_field(idx);
    return f < 0 ? nullptr : s.E + (size_t)f * s.V;
}

// (Re)initialise a slot for parameter set p: tables, argument blocks, zero
// state, q[0] = 1 (SCFT.c:785-843, 881-981).
int slot_init(scft_ctx *c, Slot &s, const scft_params &p)
{
    scft_derived d;
    int rc = scft_derive(&p, &d);
    if (rc != SCFT_OK) {
        set_err(rc == SCFT_EVOLUME ? "volumes do not sum to 1" : "invalid parameters");
        return rc;
    }
    if (!same_structure(p, c->proto) || d.ny != c->NY || d.nz != c->NZ) {
        set_err("slot parameters do not share the batch's chain structure");
        return SCFT_EINVAL;
    }
    if (d.nx > c->max_nx) {
        set_err("slot NX exceeds the batch capacity");
        return SCFT_EINVAL;
    }
    s.p = p;
    s.d = d;
    s.NX = d.nx;
    s.NY = d.ny;
    s.NZ = d.nz;
    s.V = d.nx * d.ny * d.nz;
    const size_t V = s.V;
    const double DS = 1.0 / (p.ns - 1);
    const double DX = p.lx / s.NX, DY = p.ly / s.NY, DZ = p.lz / s.NZ;

    // transform tables (cached per context)
    auto find_fft = [&](int L, double alpha, const FftTable **out) -> int {
        for (auto &f : c->ffts) {
            if (f->L != L) continue;
            // same K? check one entry on the host copy by recomputation
            double kh = axis_k(1 % L, L, alpha), kd = 0.0;
            HIPCHK(hipMemcpy(&kd, f->K + (1 % L), sizeof(double), hipMemcpyDeviceToHost));
            if (kh == kd) { *out = f.get(); return SCFT_OK; }
        }
        return get_fft(c, L, alpha, out);
    };
    const int dir = p.mask ? 1 : 0;
    s.fz = nullptr;
    s.dz = nullptr;
    const double az = DS / (DZ * DZ);
    if (c->zkind == Z_DENSE) {
        for (auto &t : c->denses)
            if (t->n == s.NZ) s.dz = t.get();
        if (!s.dz) RETCHK(get_dense(c, s.NZ, dir, az, &s.dz));
    } else {
        RETCHK(find_fft(dir ? 2 * (s.NZ + 1) : s.NZ, az, &s.fz));
    }
    s.fy = nullptr;
    if (s.NY > 1) RETCHK(find_fft(s.NY, DS / (DY * DY), &s.fy));
    s.fx = nullptr;
    if (s.NX > 1) RETCHK(find_fft(s.NX, DS / (DX * DX), &s.fx));

    // zero state
    fill_d(c, s.w, 0.0, 13 * V);
    fill_d(c, s.pf, 0.0, V);
    fill_d(c, s.tot, 0.0, 13 * V);
    fill_d(c, s.phitot, 0.0, V);
    fill_d(c, s.mask, 0.0, 3 * V);
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        if (d.ns_sp[sp] <= 0) continue;
        fill_t(c, s.q[sp], 1.0, V);
        if (s.qd[sp] != s.q[sp]) fill_t(c, s.qd[sp], 1.0, V);
    }

    // prep
    PrepArgs &pa = s.prep;
    std::memset(&pa, 0, sizeof(pa));
    bool used[13] = {false};
    for (auto &ch : s.chains) used[ch.fa] = used[ch.fb] = true;
    for (int f = 0; f < 13; ++f)
        if (used[f]) pa.fields[pa.nfields++] = f;
    pa.w = s.w;
    pa.E = s.E;
    pa.E2f = s.E2f;
    pa.q0_dbk1 = d.ns_sp[0] > 0 ? s.q[0] : nullptr;
    pa.ds = DS;
    pa.V = s.V;
    pa.nx = s.NX;
    pa.ny = s.NY;
    pa.nz = s.NZ;
    pa.delta = p.seed_area_delta;

    // Q slots: dbk1/2 from q-dagger, dbk3/4 and the rest from q
    // (SCFT.c:1271, 1314, 1357, 1399, 1422); ln Q coefficients (SCFT.c:1529-1555)
    FinArgs &fin = s.fin;
    std::memset(&fin, 0, sizeof(fin));
    s.qslots.clear();
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        const int ns = d.ns_sp[sp];
        if (ns <= 0) continue;
        const void *src = (sp < 2) ? s.qd[sp] : s.q[sp];
        const Chain *ch = find_chain(s, src);
        if (!ch) {
            set_err("internal: no chain for a Q slice");
            return SCFT_EINVAL;
        }
        QSlot qs;
        qs.end = static_cast<const char *>(src) + (size_t)(ns - 1) * V * c->tsize;
        qs.e = slice_factor(s, *ch, ns - 1);
        qs.partial = s.qpart + (size_t)fin.nq * c->nred;   // k_qfinal reads [nq][nred]
        qs.V = s.V;
        s.qslots.push_back(qs);
        fin.qsp[fin.nq] = sp;
        double coef;
        if (sp < 4) coef = p.vol_dbk[sp] / d.alpha[sp] * (1.0 - d.coating_vol);
        else if (sp < 14) coef = p.vol_homo[sp - 4] / d.alpha[sp] * (1.0 - d.coating_vol);
        else coef = p.mask ? p.vol_s[sp - 14] / d.alpha[sp] : 0.0;
        fin.lnq_coef[fin.nq] = coef;
        fin.nq++;
    }
    fin.V = s.V;
    fin.nblocks = c->nred;
    fin.nact = d.n_active;
    fin.qpart = s.qpart;
    fin.dpart = s.dpart;
    fin.upart = s.upart;
    fin.p = s.pf;
    fin.out = c->scal + (&s - c->slots.data());

    // density terms (SCFT.c:1092-1513)
    DenArgs &den = s.den;
    std::memset(&den, 0, sizeof(den));
    auto term = [&](int sp, int start, int stop, int field, double frac) {
        DenTerm &t = den.term[den.nterms++];
        t.q = s.q[sp];
        t.qd = s.qd[sp];
        t.ns = d.ns_sp[sp];
        t.start = start;
        t.stop = stop;
        t.field = field;
        t.sp = sp;
        t.frac = frac;
        const Chain *cq = find_chain(s, t.q), *cd = find_chain(s, t.qd);
        t.eqa = s.E + (size_t)cq->fa * V;
        t.eqb = s.E + (size_t)cq->fb * V;
        t.qsw = cq->switch_at;
        t.eda = s.E + (size_t)cd->fa * V;
        t.edb = s.E + (size_t)cd->fb * V;
        t.dsw = cd->switch_at;
    };
    const double cv = d.coating_vol;
    for (int h = 0; h < 10; ++h)
        if (d.ns_sp[4 + h] > 0) term(4 + h, 0, d.ns_sp[4 + h] - 1, h, p.vol_homo[h] * (1.0 - cv));
    for (int b = 0; b < 4; ++b) {
        if (d.ns_sp[b] <= 0) continue;
        const double fr = p.vol_dbk[b] * (1.0 - cv);
        term(b, 0, d.ns_x[b] - 1, kDbkField[b][0], fr);
        term(b, d.ns_x[b] - 1, d.ns_sp[b] - 1, kDbkField[b][1], fr);
    }
    for (int b = 0; b < 3; ++b)
        if (d.ns_sp[14 + b] > 0) term(14 + b, 0, d.ns_sp[14 + b] - 1, 10 + b, p.mask ? p.vol_s[b] : 0.0);
    bool carries[13] = {false};
    for (int i = 0; i < den.nterms; ++i) carries[den.term[i].field] = true;
    for (int f = 0; f < 13; ++f)
        if (carries[f]) den.phif[den.nphi++] = f;
    for (int f = 0; f < 13; ++f)
        for (int g = 0; g < 13; ++g) den.chi[f][g] = d.chi_eff[f][g];
    den.V = s.V;
    den.mask = p.mask;
    den.scal = fin.out;
    den.w = s.w;
    den.p = s.pf;
    den.tot = s.tot;
    den.phitot = s.phitot;
    den.partial = s.dpart;

    // update (SCFT.c:1993-2124)
    UpdArgs &u = s.upd;
    std::memset(&u, 0, sizeof(u));
    u.V = s.V;
    u.nx = s.NX;
    u.ny = s.NY;
    u.nz = s.NZ;
    for (int f = 0; f < 13; ++f)
        if (d.active[f]) u.act[u.nact++] = f;
    for (int f = 0; f < 13; ++f) {
        u.phibar[f] = d.phi_avg[f];
        for (int g = 0; g < 13; ++g) u.chi[f][g] = d.chi_eff[f][g];
    }
    u.g1 = p.g1;
    u.g2 = p.g1 / 6.0;
    u.w = s.w;
    u.p = s.pf;
    u.tot = s.tot;
    u.partial = s.upart;
    s.ready = true;
    c->dirty = true;
    return SCFT_OK;
}

// ------------------------------------------------------------- schedule

template <typename X>
int upload(scft_ctx *c, X **dev, size_t *cap, const std::vector<X> &v)
{
    if (v.size() > *cap) {
        if (*dev) {
            auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)*dev);
            if (it != c->allocs.end()) c->allocs.erase(it);
            (void)hipFree(*dev);
        }
        size_t n = std::max<size_t>(v.size(), 2 * *cap);
        RETCHK(ctx_alloc(c, dev, n));
        *cap = n;
    }
    if (!v.empty())
        HIPCHK(hipMemcpyAsync(*dev, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, c->stream));
    return SCFT_OK;
}

void drop_graph(scft_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// Key of the unbiased rounding of the FP32 slice a launch stores: chain ci
// of its slot, contour step t (the same for a slot wherever it sits in a
// batch); 0 for launches that store the scratch between axes, and when the
// context rounds to nearest.
unsigned round_key(const scft_ctx *c, int ci, int t, int ai, bool last)
{
    if (!c->sround || c->tsize != 4 || !last) return 0;
    const unsigned k = (0x9E3779B1u * unsigned(ci + 1)) ^ (0x85EBCA6Bu * unsigned(t + 1)) ^ (0xC2B2AE35u * unsigned(ai + 1));
    return k ? k : 1u;
}

// x lengths of the fused film step (the register-FFT forms of 16 lanes)
bool film_length(int L) { return L == 192 || L == 256 || L == 320; }

// periodic lengths of k_fftx (32-line LDS groups)
bool fftx_length(int L) { return L == 128 || L == 192 || L == 256 || L == 320; }

// The film step against the split step is tuned for three sizes of the
// active set: it wins where few chains share a launch (its fused launch
// halves the launches and keeps the scratch on chip) and can lose on a full
// batch (each part re-stages whole rows for a quarter of the products).
constexpr int kFilmTuneSlots[3] = {8, 32, 128};
int film_class(int nact) { return nact <= 16 ? 0 : (nact <= 64 ? 1 : 2); }

// The context's kernels allow the fused film step: 2D (NY = 1), NZ = 128
// on the even MFMA dense operator, the scratch in the slice type.
bool film_fits(const scft_ctx *c)
{
    return c->NY == 1 && c->NZ == kFilmN && c->dense_mfma && !c->dense_mfma1 && !c->tmp_wide && c->zkind == Z_DENSE;
}

// The register-FFT launches of one step and axis for lengths 192, 256 and
// 320 (from launches[lfirst] on, their ops consecutive) become one
// LK_FFTMIX launch: the workgroups of two per CU are split between the
// lengths in proportion to units x length / waves per workgroup, each group
// at least one workgroup and at most one unit per wave.  FP32 slices and
// scratch only: with FP64 lines the 320-point body needs more than 256
// registers (one wave per SIMD), and the mixed kernel would hold the other
// two lengths to that.
int merge_mix(scft_ctx *c, size_t lfirst)
{
    std::vector<Launch> &Ls = c->launches;
    const size_t n = Ls.size() - lfirst;
    if (n < 2 || c->tsize != 4 || c->tmp_wide) return SCFT_OK;
    auto group = [](const Launch &ln) {
        if (ln.kind == LK_FFTREG && ln.L == 256) return 0;
        if (ln.kind == LK_FFTREG2 && ln.L == 192) return 1;
        if (ln.kind == LK_FFTREG2 && ln.L == 320) return 2;
        return -1;
    };
    const int waves[3] = {kFRWaves, FReg2<12>::W, FReg2<20>::W};
    double w[3] = {0.0, 0.0, 0.0}, tot = 0.0;
    int units[3] = {0, 0, 0};
    Launch m{};
    m.kind = LK_FFTMIX;
    m.pos = Ls[lfirst].pos;
    m.first = Ls[lfirst].first;
    m.count = 0;
    for (size_t i = lfirst; i < Ls.size(); ++i) {
        const int k = group(Ls[i]);
        if (k < 0 || Ls[i].first != m.first + m.count) return SCFT_OK;   // not a mix of these lengths
        MixGroup &gr = m.mix.g[k];
        gr.op0 = Ls[i].first - m.first;
        gr.upo = Ls[i].P;
        gr.nunits = Ls[i].count * Ls[i].P;
        units[k] = gr.nunits;
        w[k] = double(gr.nunits) * Ls[i].L / waves[k];
        tot += w[k];
        m.count += Ls[i].count;
        m.L = std::max(m.L, Ls[i].L);
    }
    int wg = 0;
    for (int k = 0; k < 3; ++k) {
        MixGroup &gr = m.mix.g[k];
        gr.wg0 = wg;
        if (units[k] == 0) continue;
        const int most = (units[k] + waves[k] - 1) / waves[k];
        gr.nwg = std::max(1, std::min(most, int(std::lround(2.0 * c->ncu * w[k] / tot))));
        wg += gr.nwg;
    }
    m.gx = wg;
    Ls.resize(lfirst);
    Ls.push_back(m);
    return SCFT_OK;
}

int build_schedule(scft_ctx *c)
{
    drop_graph(c);
    c->act.clear();
    for (int i = 0; i < c->nslots; ++i)
        if (c->slots[i].ready && c->slots[i].active) c->act.push_back(i);
    c->launches.clear();
    c->nstep0 = 0;
    c->step_first.clear();
    std::vector<LineOp> lops;
    std::vector<DenseOp> dops;
    std::vector<PlaneOp> pops;
    std::vector<FilmOp> fops;
    int maxsteps = 0;
    for (int si : c->act)
        for (auto &ch : c->slots[si].chains) maxsteps = std::max(maxsteps, ch.nsteps);
    const size_t ts = c->tsize;
    // every active slot a periodic 64 x 64 plane: one fused launch per step
    bool plane = c->plane_on && c->fft_reg && c->zkind == Z_FFT && c->NY == 1 && c->NZ == kPlaneN;
    for (int si : c->act) plane &= c->slots[si].NX == kPlaneN && !c->slots[si].p.mask;
    // every active slot a 2D film with the dense z operator of NZ = 128 and
    // a register-FFT x length: one fused launch per step and x length
    bool film = film_fits(c) && !c->act.empty() && c->film_cls[film_class(int(c->act.size()))];
    std::vector<int> film_lens;
    for (int si : c->act) {
        const Slot &s = c->slots[si];
        film &= film_length(s.NX) && s.dz && s.dz->hs == kFilmN / 2 && s.dz->ha == kFilmN / 2 &&
                s.dz->MsT == c->slots[c->act[0]].dz->MsT;
        if (std::find(film_lens.begin(), film_lens.end(), s.NX) == film_lens.end()) film_lens.push_back(s.NX);
    }
    for (int t = 0; t < maxsteps; ++t) {
        c->step_first.push_back(int(c->launches.size()));
        if (film) {
            for (int L : film_lens) {
                Launch ln{};
                ln.kind = LK_FILM;
                ln.pos = 3;
                ln.L = L;
                ln.first = int(fops.size());
                const Slot *s0 = nullptr;
                for (int si : c->act) {
                    Slot &s = c->slots[si];
                    if (s.NX != L) continue;
                    if (!s0) s0 = &s;
                    const size_t V = s.V;
                    for (size_t ci = 0; ci < s.chains.size(); ++ci) {
                        const Chain &ch = s.chains[ci];
                        if (ch.nsteps <= t) continue;
                        FilmOp op{};
                        op.src = static_cast<const char *>(ch.base) + (size_t)t * V * ts;
                        op.dst = static_cast<char *>(ch.base) + (size_t)(t + 1) * V * ts;
                        op.epre = s.E + (size_t)ch.step_field(t) * V;
                        op.epre2 = slice_factor(s, ch, t);
                        op.e2f = (s.E2f && op.epre2 == op.epre) ? s.E2f + (size_t)ch.step_field(t) * V : nullptr;
                        op.Kx = s.fx->K;
                        // the key of the split step's x launch (axis 1 of z, x)
                        op.rkey = round_key(c, int(ci), t, 1, true);
                        fops.push_back(op);
                    }
                }
                ln.count = int(fops.size()) - ln.first;
                if (ln.count == 0) continue;
                ln.ftab.MsT = s0->dz->MsT;
                ln.ftab.MaT = s0->dz->MaT;
                ln.ftab.ldm = s0->dz->ldm;
                ln.ftab.tw = reinterpret_cast<const double *>(s0->fx->tw);
                ln.gx = 32 * ((ln.count + 7) / 8);   // kFilmParts parts of every op, XCD-grouped
                c->launches.push_back(ln);
            }
            if (t == 0) c->nstep0 = int(c->launches.size());
            continue;
        }
        if (plane) {
            Launch ln{};
            ln.kind = LK_PLANE;
            ln.pos = 3;
            ln.L = kPlaneN;
            ln.first = int(pops.size());
            for (int si : c->act) {
                Slot &s = c->slots[si];
                const size_t V = s.V;
                for (size_t ci = 0; ci < s.chains.size(); ++ci) {
                    const Chain &ch = s.chains[ci];
                    if (ch.nsteps <= t) continue;
                    PlaneOp op{};
                    op.rkey = round_key(c, int(ci), t, 0, true);
                    op.src = static_cast<const char *>(ch.base) + (size_t)t * V * ts;
                    op.dst = static_cast<char *>(ch.base) + (size_t)(t + 1) * V * ts;
                    op.epre = s.E + (size_t)ch.step_field(t) * V;
                    op.epre2 = slice_factor(s, ch, t);
                    op.Kz = s.fz->K;
                    op.Kx = s.fx->K;
                    op.twz = reinterpret_cast<const double *>(s.fz->tw);
                    op.twx = reinterpret_cast<const double *>(s.fx->tw);
                    pops.push_back(op);
                }
            }
            ln.count = int(pops.size()) - ln.first;
            // persistent workgroups (36 KB of LDS each: four per CU)
            ln.gx = std::max(1, std::min(ln.count, 4 * c->ncu));
            if (ln.count > 0) c->launches.push_back(ln);
            if (t == 0) c->nstep0 = int(c->launches.size());
            continue;
        }
        // axes in order z, y, x; every slot of the batch shares NY/NZ, so the
        // axis list is the same for all (x exists when any slot has NX > 1)
        std::vector<int> axes = {2};
        if (c->NY > 1) axes.push_back(1);
        bool anyx = false;
        for (int si : c->act) anyx |= c->slots[si].NX > 1;
        if (anyx) axes.push_back(0);
        const int nax = int(axes.size());
        for (int ai = 0; ai < nax; ++ai) {
            const int axis = axes[ai];
            const int pos = (nax == 1) ? 3 : (ai == 0 ? 0 : (ai == nax - 1 ? 2 : 1));
            const bool dense = (axis == 2 && c->zkind == Z_DENSE);
            std::vector<LineOp> tl;     // this step/axis, every live chain
            std::vector<DenseOp> td;
            for (int si : c->act) {
                Slot &s = c->slots[si];
                const size_t V = s.V;
                for (size_t ci = 0; ci < s.chains.size(); ++ci) {
                    const Chain &ch = s.chains[ci];
                    if (ch.nsteps <= t) continue;
                    const unsigned rkey = round_key(c, int(ci), t, ai, ai == nax - 1);
                    const void *src = (ai == 0) ? static_cast<const char *>(ch.base) + (size_t)t * V * ts : (const void *)ch.tmp;
                    void *dst = (ai == nax - 1) ? static_cast<char *>(ch.base) + (size_t)(t + 1) * V * ts : (void *)ch.tmp;
                    // first axis: E of this step times the factor the stored slice owes
                    const double *ep = (ai == 0) ? s.E + (size_t)ch.step_field(t) * V : nullptr;
                    const double *ep2 = (ai == 0) ? slice_factor(s, ch, t) : nullptr;
                    if (dense) {
                        DenseOp op{};
                        op.src = src;
                        op.dst = dst;
                        op.epre = ep;
                        op.epre2 = ep2;
                        op.epost = nullptr;
                        // inside a block the step's factor is E^2 of one field
                        op.e2f = (s.E2f && ep && ep2 == ep) ? s.E2f + (size_t)ch.step_field(t) * V : nullptr;
                        op.MsT = s.dz->MsT;
                        op.MaT = s.dz->MaT;
                        op.n = s.dz->n;
                        op.hs = s.dz->hs;
                        op.ha = s.dz->ha;
                        op.ldm = s.dz->ldm;
                        op.nrows = s.NX * s.NY;
                        td.push_back(op);
                        continue;
                    }
                    const FftTable *ft = axis == 2 ? s.fz : (axis == 1 ? s.fy : s.fx);
                    LineOp op{};
                    op.src = src;
                    op.dst = dst;
                    op.epre = ep;
                    op.epre2 = ep2;
                    op.epost = nullptr;
                    op.rkey = rkey;
                    op.K = ft->K;
                    op.tw = ft->tw;
                    op.L = ft->L;
                    op.npass = ft->rad.npass;
                    for (int i = 0; i < kMaxPass; ++i) op.radix[i] = i < ft->rad.npass ? ft->rad.r[i] : 1;
                    if (axis == 2) {
                        op.n = s.NZ;
                        op.dirichlet = s.p.mask ? 1 : 0;
                        op.nlines = s.NX * s.NY;
                        op.c1 = 1; op.s0 = 0; op.s1 = s.NZ; op.stride = 1;
                    } else if (axis == 1) {
                        op.n = s.NY;
                        op.nlines = s.NX * s.NZ;
                        op.c1 = s.NZ; op.s0 = 1; op.s1 = s.NZ * s.NY; op.stride = s.NZ;
                    } else {
                        op.n = s.NX;
                        op.nlines = s.NY * s.NZ;
                        op.c1 = s.NY * s.NZ; op.s0 = 1; op.s1 = 0; op.stride = s.NY * s.NZ;
                    }
                    // strided lines 2i, 2i+1 sit next to each other at an even offset
                    op.pairs = (op.s0 == 1 && op.c1 % 2 == 0 && op.s1 % 2 == 0 && op.stride % 2 == 0) ? 1 : 0;
                    tl.push_back(op);
                }
            }
            if (dense && !td.empty()) {
                Launch ln{};
                ln.pos = pos;
                ln.first = int(dops.size());
                ln.count = int(td.size());
                int maxl = 0;
                for (auto &op : td) maxl = std::max(maxl, op.nrows);
                const int hsp = ((td[0].hs + 15) / 16) * 16;
                // the MFMA kernels take E on load (first axis) and no factor on store
                bool one_table = true;
                for (auto &op : td) one_table &= op.MsT == td[0].MsT && op.n == td[0].n && op.epre && !op.epost;
                if (c->dense_mfma && hsp <= 64 && one_table) {
                    // persistent waves over (op, 16-row tile); one workgroup per CU
                    const bool even = td[0].hs == td[0].ha && !c->dense_mfma1;
                    const int wpg = even ? kMfma2Waves : kMfmaWaves;
                    // the prefetching form takes FP32 rows of NZ 128, 64 or 48 in whole 16-row
                    // tiles and the FP64 factor
                    const int n0 = td[0].n;
                    bool pf = even && c->zpf && ts == 4 && !c->tmp_wide &&
                              ((hsp == 64 && n0 == 128) || (hsp == 32 && (n0 == 64 || n0 == 48)));
                    for (auto &op : td) pf &= op.n == n0 && op.nrows % 16 == 0 && !op.e2f;
                    ln.pf = pf;
                    if (pf) ln.L = n0;   // the row length of the prefetching form
                    ln.kind = even ? LK_DENSE_MFMA2 : LK_DENSE_MFMA;
                    ln.hsp = hsp;
                    ln.P = (maxl + 15) / 16;   // tiles per op
                    int waves = ln.count * ln.P;
                    if (even) {
                        // ops grouped by row count (uneven NX), each with its own tiles
                        std::stable_sort(td.begin(), td.end(),
                                         [](const DenseOp &a, const DenseOp &b) { return a.nrows > b.nrows; });
                        ZTiles zt{};
                        int nt = 0;
                        for (size_t i = 0; i < td.size();) {
                            size_t j = i;
                            while (j < td.size() && td[j].nrows == td[i].nrows) ++j;
                            if (zt.ng == kZGroups) {   // more row counts than groups: one tile count for all
                                zt = ZTiles{};
                                zt.ng = 1;
                                zt.tpo[0] = ln.P;
                                nt = ln.count * ln.P;
                                break;
                            }
                            zt.op0[zt.ng] = int(i);
                            zt.tpo[zt.ng] = (td[i].nrows + 15) / 16;
                            zt.t0[zt.ng] = nt;
                            nt += int(j - i) * zt.tpo[zt.ng];
                            ++zt.ng;
                            i = j;
                        }
                        ln.zt = zt;
                        ln.ntiles = nt;
                        waves = nt;
                    }
                    ln.gx = std::max(1, std::min((waves + wpg - 1) / wpg, c->ncu));
                } else {
                    ln.kind = LK_DENSE;
                    ln.gx = (maxl + kDenseTM - 1) / kDenseTM;
                }
                dops.insert(dops.end(), td.begin(), td.end());
                c->launches.push_back(ln);
            }
            // one launch per transform length (slots of a batch may differ in NX)
            const size_t lfirst = c->launches.size();
            std::vector<int> Ls;
            for (auto &op : tl)
                if (std::find(Ls.begin(), Ls.end(), op.L) == Ls.end()) Ls.push_back(op.L);
            for (int L : Ls) {
                Launch ln{};
                ln.pos = pos;
                ln.first = int(lops.size());
                int maxl = 0;
                bool dirichlet = false;
                for (auto &op : tl)
                    if (op.L == L) {
                        lops.push_back(op);
                        maxl = std::max(maxl, op.nlines);
                        dirichlet |= op.dirichlet != 0;
                    }
                ln.count = int(lops.size()) - ln.first;
                ln.L = L;
                const bool reg1 = c->fft_reg && !dirichlet && (L == 128 || L == 256);
                const bool reg2 = c->fft_reg && !dirichlet && (L == 192 || L == 320 || L == 96 || L == 64);
                const int unit = reg1 ? (L == 256 ? FReg<16>::UNIT : FReg<8>::UNIT)
                                      : (L <= 96 ? FReg2<12, 8>::UNIT : FReg2<12>::UNIT);
                const int units = ln.count * ((maxl + unit - 1) / unit);
                // A register-FFT launch of fewer units (one wave each) than two
                // per CU leaves most SIMDs idle for the whole transform; the
                // wave-per-line FFT gives every line pair its own wave there.
                const bool few = !c->fft_reg_force && c->fft_wave && fftwave_supported(L) && units < 2 * c->ncu;
                // 32 lines that start at 32 consecutive elements, stored (last axis): k_fftx
                // where the autotuner measured it faster for the active set's size class
                // (on C3, 128 x 128 x 64, 4.5% faster per iteration than the 8 x 16 register
                // FFT with one solve, 4.6% slower with four)
                bool coal = !c->tmp_wide && !dirichlet && fftx_length(L) &&
                            (c->fftx_mode == 1 || (c->fftx_mode == 2 && c->xcoal_cls[film_class(int(c->act.size()))]));
                for (int k = ln.first; coal && k < ln.first + ln.count; ++k) {
                    const LineOp &op = lops[k];
                    coal = !op.epre && !op.epost && op.s0 == 1 && op.c1 % 32 == 0 && op.s1 % 32 == 0 &&
                           op.stride % 32 == 0 && op.nlines % 32 == 0;
                }
                if (coal) {
                    ln.kind = LK_FFTX;
                    ln.P = (maxl + 31) / 32;   // 32-line groups per op
                    ln.gx = ln.count * ln.P;
                } else if (reg1 && !few) {
                    // persistent waves over (op, unit of 2G real lines); two workgroups per CU
                    ln.kind = LK_FFTREG;
                    ln.P = (maxl + unit - 1) / unit;   // units per op
                    ln.gx = std::max(1, std::min((units + kFRWaves - 1) / kFRWaves, 2 * c->ncu));
                } else if (reg2 && !few) {
                    // the same over 16 lanes of 12 or 20 points, 8 lanes of 12 (96) or of 8 (64)
                    ln.kind = LK_FFTREG2;
                    const int wpg = (L == 320) ? FReg2<20>::W : FReg2<12>::W;
                    const int wgcu = (L == 64) ? FReg2<8, 8>::WG_PER_CU : FReg2<12>::WG_PER_CU;
                    ln.P = (maxl + unit - 1) / unit;
                    ln.gx = std::max(1, std::min((units + wpg - 1) / wpg, wgcu * c->ncu));
                } else if (c->fft_wave && !dirichlet && fftwave_supported(L)) {
                    ln.kind = LK_FFTWAVE;
                    ln.gx = (maxl + 2 * kFftWaves - 1) / (2 * kFftWaves);
                } else {
                    ln.kind = LK_LINEOP;
                    ln.P = lines_P(L);
                    ln.lds = size_t(2) * ln.P * L * sizeof(cpx<double>);
                    ln.gx = (maxl + 2 * ln.P - 1) / (2 * ln.P);
                }
                c->launches.push_back(ln);
            }
            if (c->xmix) RETCHK(merge_mix(c, lfirst));
        }
        if (t == 0) c->nstep0 = int(c->launches.size());
    }
    for (size_t t = 0; t < c->step_first.size(); ++t) {
        const size_t e = t + 1 < c->step_first.size() ? size_t(c->step_first[t + 1]) : c->launches.size();
        for (size_t i = size_t(c->step_first[t]); i < e; ++i) c->launches[i].step = int(t);
    }
    // side streams for launches of one step and axis position (one per x length
    // of a batch of uneven NX, up to three), created when a schedule first has
    // such launches: every stream of the process shares its few hardware queues
    if (c->par_axes && c->side.size() < 3) {
        size_t need = 0;
        for (size_t i = 0; i < c->launches.size();) {
            size_t j = i + 1;
            while (j < c->launches.size() && c->launches[j].step == c->launches[i].step &&
                   c->launches[j].pos == c->launches[i].pos)
                ++j;
            need = std::max(need, std::min<size_t>(j - i - 1, 3));
            i = j;
        }
        while (c->side.size() < need) {
            hipStream_t st = nullptr;
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            c->side.push_back(st);
        }
    }
    RETCHK(upload(c, &c->lineops, &c->lineops_cap, lops));
    RETCHK(upload(c, &c->denseops, &c->denseops_cap, dops));
    RETCHK(upload(c, &c->planeops, &c->planeops_cap, pops));
    RETCHK(upload(c, &c->filmops, &c->filmops_cap, fops));
    // per-slot argument arrays in active order
    std::vector<PrepArgs> prep;
    std::vector<DenArgs> den;
    std::vector<UpdArgs> upd;
    std::vector<FinArgs> fin;
    std::vector<QSlot> qs;
    for (int si : c->act) {
        Slot &s = c->slots[si];
        prep.push_back(s.prep);
        den.push_back(s.den);
        upd.push_back(s.upd);
        fin.push_back(s.fin);
        for (auto &q : s.qslots) qs.push_back(q);
    }
    size_t cap = c->slots.size();
    if (!c->d_prep) {
        RETCHK(ctx_alloc(c, &c->d_prep, cap));
        RETCHK(ctx_alloc(c, &c->d_den, cap));
        RETCHK(ctx_alloc(c, &c->d_upd, cap));
        RETCHK(ctx_alloc(c, &c->d_fin, cap));
    }
    if (!prep.empty()) {
        HIPCHK(hipMemcpyAsync(c->d_prep, prep.data(), prep.size() * sizeof(PrepArgs), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_den, den.data(), den.size() * sizeof(DenArgs), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_upd, upd.data(), upd.size() * sizeof(UpdArgs), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_fin, fin.data(), fin.size() * sizeof(FinArgs), hipMemcpyHostToDevice, c->stream));
    }
    RETCHK(upload(c, &c->d_qslots, &c->d_qslots_cap, qs));
    c->nqslots = int(qs.size());
    std::vector<RngSlot> rs;
    c->rng_maxchunk = 0;
    for (int si : c->act) {
        Slot &s = c->slots[si];
        if (!s.rng_on || s.rng_need == 0) continue;
        rs.push_back(RngSlot{c->rng + si, s.rng_counts, s.rng_offsets, s.noise, s.rng_need, s.rng_nchunk});
        c->rng_maxchunk = std::max(c->rng_maxchunk, s.rng_nchunk);
    }
    RETCHK(upload(c, &c->d_rng, &c->d_rng_cap, rs));
    c->nrng = int(rs.size());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->dirty = false;
    return SCFT_OK;
}

// Per-iteration argument changes (seed mode, noise) go to the device
// argument blocks; the launch sequence itself does not change.
int set_iteration_args(scft_ctx *c, float mf, const double *noise_dev)
{
    bool changed = false;
    for (size_t a = 0; a < c->act.size(); ++a) {
        Slot &s = c->slots[c->act[a]];
        int mode = 0;
        if (s.d.ns_sp[0] > 0 && s.p.enable_seed) mode = (s.iter <= (unsigned long)s.p.seed_steps) ? 1 : 2;
        if (mode != s.prep.seed_mode) {
            s.prep.seed_mode = mode;
            HIPCHK(hipMemcpyAsync(c->d_prep + a, &s.prep, sizeof(PrepArgs), hipMemcpyHostToDevice, c->stream));
            changed = true;
        }
        // mf < 0: each slot's own MF (1 while its iteration is within RELAXATION)
        const float smf = mf >= 0.0f ? mf : ((s.rng_on && s.iter <= s.p.relaxation) ? 1.0f : 0.0f);
        const double amp = s.p.dt_noise * smf;
        const double *nz = (smf != 0.0f) ? (mf >= 0.0f ? noise_dev : s.noise) : nullptr;
        if (amp != s.upd.amp || nz != s.upd.noise) {
            s.upd.amp = amp;
            s.upd.noise = nz;
            HIPCHK(hipMemcpyAsync(c->d_upd + a, &s.upd, sizeof(UpdArgs), hipMemcpyHostToDevice, c->stream));
            changed = true;
        }
    }
    (void)changed;
    return SCFT_OK;
}

template <typename T>
void launch_prep(scft_ctx *c, int nact, int maxV)
{
    hipLaunchKernelGGL(k_prep<T>, dim3(grid_for(maxV, 512), nact), dim3(kThreads), 0, c->stream,
                       (const PrepArgs *)c->d_prep);
}

// Typed launchers: slices are T; the scratch between axes is T too, or FP64
// when tmp_wide (then the axis position fixes the input/output types).
template <typename Tin, typename Tout>
void launch_fftwave(scft_ctx *c, const Launch &ln, hipStream_t st)
{
    const LineOp *ops = c->lineops + ln.first;
    dim3 g(ln.gx, ln.count), b(64 * kFftWaves);
    switch (ln.L) {
#define SCFT_FFTW_CASE(LL) \
    case LL: hipLaunchKernelGGL((k_fftwave<LL, Tin, Tout>), g, b, 0, st, ops); break;
    SCFT_FFTW_CASE(16) SCFT_FFTW_CASE(32) SCFT_FFTW_CASE(48) SCFT_FFTW_CASE(64) SCFT_FFTW_CASE(80)
    SCFT_FFTW_CASE(96) SCFT_FFTW_CASE(128) SCFT_FFTW_CASE(160) SCFT_FFTW_CASE(192) SCFT_FFTW_CASE(256)
    SCFT_FFTW_CASE(320) SCFT_FFTW_CASE(384) SCFT_FFTW_CASE(512)
#undef SCFT_FFTW_CASE
    default: break;
    }
}

template <typename Tin, typename Tout>
void launch_one(scft_ctx *c, const Launch &ln, hipStream_t st)
{
    dim3 g(ln.gx, ln.count);
    switch (ln.kind) {
    case LK_PLANE:
        hipLaunchKernelGGL((k_plane64<Tin, Tout>), dim3(ln.gx), dim3(64 * kPlaneWaves), 0, st,
                           c->planeops + ln.first, ln.count);
        break;
    case LK_FILM: {
        // slices and the on-chip scratch are T (the film step is never scheduled with a wide scratch)
        const FilmOp *ops = c->filmops + ln.first;
        const dim3 g1(ln.gx), b(64 * kFilmWaves);
        if (ln.L == 256) hipLaunchKernelGGL((k_film<256, Tin>), g1, b, 0, st, ops, ln.count, ln.ftab);
        else if (ln.L == 192) hipLaunchKernelGGL((k_film<192, Tin>), g1, b, 0, st, ops, ln.count, ln.ftab);
        else hipLaunchKernelGGL((k_film<320, Tin>), g1, b, 0, st, ops, ln.count, ln.ftab);
        break;
    }
    case LK_DENSE_MFMA: {
        const DenseOp *ops = c->denseops + ln.first;
        const dim3 g1(ln.gx), b(64 * kMfmaWaves);
        const int nt = ln.count * ln.P;
        if (ln.hsp == 16) hipLaunchKernelGGL((k_zdense_mfma<Tin, Tout, 16>), g1, b, 0, st, ops, nt, ln.P);
        else if (ln.hsp == 32) hipLaunchKernelGGL((k_zdense_mfma<Tin, Tout, 32>), g1, b, 0, st, ops, nt, ln.P);
        else if (ln.hsp == 48) hipLaunchKernelGGL((k_zdense_mfma<Tin, Tout, 48>), g1, b, 0, st, ops, nt, ln.P);
        else hipLaunchKernelGGL((k_zdense_mfma<Tin, Tout, 64>), g1, b, 0, st, ops, nt, ln.P);
        break;
    }
    case LK_DENSE_MFMA2: {
        const DenseOp *ops = c->denseops + ln.first;
        const dim3 g1(ln.gx), b(64 * kMfma2Waves);
        const int nt = ln.ntiles;
        const ZTiles &zt = ln.zt;
        // the grouped tile map only where row counts differ (it costs a
        // little on a uniform batch)
        auto go = [&](auto gr) {
            constexpr bool GR = decltype(gr)::value;
            if constexpr (sizeof(Tout) == 4) {
                if (ln.pf) {
                    if (ln.hsp == 64)
                        hipLaunchKernelGGL((k_zdense_mfma2p<Tin, 64, 128, GR>), g1, b, 0, st, ops, nt, zt);
                    else if (ln.L == 64)
                        hipLaunchKernelGGL((k_zdense_mfma2p<Tin, 32, 64, GR>), g1, b, 0, st, ops, nt, zt);
                    else
                        hipLaunchKernelGGL((k_zdense_mfma2p<Tin, 32, 48, GR>), g1, b, 0, st, ops, nt, zt);
                    return;
                }
            }
            if (ln.hsp == 16) hipLaunchKernelGGL((k_zdense_mfma2<Tin, Tout, 16, GR>), g1, b, 0, st, ops, nt, zt);
            else if (ln.hsp == 32) hipLaunchKernelGGL((k_zdense_mfma2<Tin, Tout, 32, GR>), g1, b, 0, st, ops, nt, zt);
            else if (ln.hsp == 48) hipLaunchKernelGGL((k_zdense_mfma2<Tin, Tout, 48, GR>), g1, b, 0, st, ops, nt, zt);
            else hipLaunchKernelGGL((k_zdense_mfma2<Tin, Tout, 64, GR>), g1, b, 0, st, ops, nt, zt);
        };
        if (zt.ng > 1) go(std::true_type{});
        else go(std::false_type{});
        break;
    }
    case LK_DENSE:
        hipLaunchKernelGGL((k_zdense<Tin, Tout>), g, dim3(kThreads), 0, st, c->denseops + ln.first);
        break;
    case LK_FFTWAVE:
        launch_fftwave<Tin, Tout>(c, ln, st);
        break;
    case LK_FFTREG2:
        if (ln.L == 320)
            hipLaunchKernelGGL((k_fftreg2<20, 16, Tin, Tout>), dim3(ln.gx), dim3(64 * FReg2<20>::W), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        else if (ln.L == 96)
            hipLaunchKernelGGL((k_fftreg2<12, 8, Tin, Tout>), dim3(ln.gx), dim3(64 * FReg2<12, 8>::W), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        else if (ln.L == 64)
            hipLaunchKernelGGL((k_fftreg2<8, 8, Tin, Tout>), dim3(ln.gx), dim3(64 * FReg2<8, 8>::W), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        else
            hipLaunchKernelGGL((k_fftreg2<12, 16, Tin, Tout>), dim3(ln.gx), dim3(64 * FReg2<12>::W), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        break;
    case LK_FFTX: {
        const LineOp *ops = c->lineops + ln.first;
        const dim3 g1(ln.gx), b(64 * kFilmWaves);
        if (ln.L == 256) hipLaunchKernelGGL((k_fftx<256, Tin>), g1, b, 0, st, ops, ln.P);
        else if (ln.L == 128) hipLaunchKernelGGL((k_fftx<128, Tin>), g1, b, 0, st, ops, ln.P);
        else if (ln.L == 192) hipLaunchKernelGGL((k_fftx<192, Tin>), g1, b, 0, st, ops, ln.P);
        else hipLaunchKernelGGL((k_fftx<320, Tin>), g1, b, 0, st, ops, ln.P);
        break;
    }
    case LK_FFTMIX:
        hipLaunchKernelGGL((k_fftmix<Tin, Tout>), dim3(ln.gx), dim3(kMixThreads), 0, st, c->lineops + ln.first,
                           ln.mix);
        break;
    case LK_FFTREG:
        if (ln.L == 256)
            hipLaunchKernelGGL((k_fftreg<16, Tin, Tout>), dim3(ln.gx), dim3(64 * kFRWaves), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        else
            hipLaunchKernelGGL((k_fftreg<8, Tin, Tout>), dim3(ln.gx), dim3(64 * kFRWaves), 0, st,
                               c->lineops + ln.first, ln.count * ln.P, ln.P);
        break;
    default:
        hipLaunchKernelGGL((k_lineop<Tin, Tout>), g, dim3(kThreads), ln.lds, st, c->lineops + ln.first, ln.P);
        break;
    }
}

template <typename T>
void launch_typed(scft_ctx *c, const Launch &ln, hipStream_t st)
{
    if (!c->tmp_wide) {
        launch_one<T, T>(c, ln, st);
        return;
    }
    switch (ln.pos) {
    case 0: launch_one<T, double>(c, ln, st); break;
    case 1: launch_one<double, double>(c, ln, st); break;
    case 2: launch_one<double, T>(c, ln, st); break;
    default: launch_one<T, T>(c, ln, st); break;
    }
}

// The schedule's launches in order.  With `par`, the launches of one step
// and axis position (a batch of uneven NX has one per x length; they touch
// different chains) run side by side: the first on the context's stream,
// the others on side streams that wait for the work before them and that
// the context's stream waits for before the next launch.  Without `par`
// (timing single launches) they run one after another.
template <typename T>
int launch_steps(scft_ctx *c, bool par = false)
{
    const std::vector<Launch> &L = c->launches;
    const size_t n = L.size();
    size_t nev = 0;
    auto next_event = [&](hipEvent_t *e) -> int {
        if (nev == c->evs.size()) {
            hipEvent_t x = nullptr;
            HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            c->evs.push_back(x);
        }
        *e = c->evs[nev++];
        return SCFT_OK;
    };
    for (size_t i = 0; i < n;) {
        size_t j = i + 1;
        if (par && c->par_axes)
            while (j < n && j - i <= c->side.size() && L[j].step == L[i].step && L[j].pos == L[i].pos) ++j;
        if (j - i == 1) {
            launch_typed<T>(c, L[i], c->stream);
            i = j;
            continue;
        }
        hipEvent_t fork;
        RETCHK(next_event(&fork));
        HIPCHK(hipEventRecord(fork, c->stream));
        for (size_t k = 1; k < j - i; ++k) HIPCHK(hipStreamWaitEvent(c->side[k - 1], fork, 0));
        launch_typed<T>(c, L[i], c->stream);
        for (size_t k = 1; k < j - i; ++k) launch_typed<T>(c, L[i + k], c->side[k - 1]);
        for (size_t k = 1; k < j - i; ++k) {
            hipEvent_t join;
            RETCHK(next_event(&join));
            HIPCHK(hipEventRecord(join, c->side[k - 1]));
            HIPCHK(hipStreamWaitEvent(c->stream, join, 0));
        }
        i = j;
    }
    return SCFT_OK;
}

template <typename T>
void launch_densities(scft_ctx *c, int nact)
{
    if (c->nqslots > 0) {
        hipLaunchKernelGGL(k_qsum<T>, dim3(c->nred, c->nqslots), dim3(kThreads), 0, c->stream,
                           (const QSlot *)c->d_qslots);
        hipLaunchKernelGGL(k_qfinal, dim3(nact), dim3(kThreads), 0, c->stream, (const FinArgs *)c->d_fin);
    }
    hipLaunchKernelGGL(k_density<T>, dim3(c->nred, nact), dim3(kThreads), 0, c->stream, (const DenArgs *)c->d_den);
}

// This iteration's deviates of every drawing slot into its noise buffer,
// in draw order.
void launch_rng(scft_ctx *c)
{
    if (c->nrng == 0) return;
    const dim3 g((c->rng_maxchunk + kThreads - 1) / kThreads, c->nrng);
    hipLaunchKernelGGL(k_rng_count, g, dim3(kThreads), 0, c->stream, c->d_rng, c->rng_jump);
    hipLaunchKernelGGL(k_rng_scan, dim3(c->nrng), dim3(1024), 0, c->stream, c->d_rng);
    hipLaunchKernelGGL(k_rng_emit, g, dim3(kThreads), 0, c->stream, c->d_rng, c->rng_jump);
    hipLaunchKernelGGL(k_rng_commit, dim3(c->nrng), dim3(64), 0, c->stream, c->d_rng);
}

void launch_update(scft_ctx *c, int nact)
{
    hipLaunchKernelGGL(k_update, dim3(c->nred, nact), dim3(kThreads), 0, c->stream, (const UpdArgs *)c->d_upd);
    hipLaunchKernelGGL(k_final, dim3(nact), dim3(kThreads), 0, c->stream, (const FinArgs *)c->d_fin);
    hipLaunchKernelGGL(k_pshift, dim3(grid_for(c->max_V, 512), nact), dim3(kThreads), 0, c->stream,
                       (const FinArgs *)c->d_fin);
}

int maxV_active(scft_ctx *c)
{
    int m = 1;
    for (int si : c->act) m = std::max(m, c->slots[si].V);
    return m;
}

// phases: 1 propagate, 2 densities, 4 update, 8 copy scalars back
int enqueue(scft_ctx *c, int phases)
{
    const int nact = int(c->act.size());
    if (nact == 0) return SCFT_OK;
    const bool f32 = c->tsize == 4;
    if (phases & 1) {
        if (f32) {
            launch_prep<float>(c, nact, maxV_active(c));
            RETCHK(launch_steps<float>(c, true));
        } else {
            launch_prep<double>(c, nact, maxV_active(c));
            RETCHK(launch_steps<double>(c, true));
        }
    }
    if (phases & 2) {
        if (f32) launch_densities<float>(c, nact);
        else launch_densities<double>(c, nact);
    }
    if (phases & 4) {
        launch_rng(c);
        launch_update(c, nact);
    }
    if (phases & 8) {
        HIPCHK(hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scalars) * c->nslots, hipMemcpyDeviceToHost, c->stream));
        if (c->nrng)
            HIPCHK(hipMemcpyAsync(c->rng_host, c->rng, sizeof(RngState) * c->nslots, hipMemcpyDeviceToHost,
                                  c->stream));
    }
    HIPCHK(hipGetLastError());
    return SCFT_OK;
}

int run_iteration(scft_ctx *c, float mf, const double *noise_dev, bool wait = true)
{
    if (c->dirty) RETCHK(build_schedule(c));
    RETCHK(set_iteration_args(c, mf, noise_dev));
    if (c->act.empty()) return SCFT_OK;
    if (c->use_graph) {
        if (!c->gexec) {
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue(c, 1 | 2 | 4 | 8);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != SCFT_OK) return rc;
            if (e != hipSuccess) {
                set_err(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                return SCFT_EHIP;
            }
            c->graph = g;
            HIPCHK(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    } else {
        RETCHK(enqueue(c, 1 | 2 | 4 | 8));
    }
    if (wait) HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

// Entry points that read or change a context refuse while a batched
// iteration started by scft_batch_launch runs on its stream: the schedule's
// argument arrays and the slots' buffers belong to that launch until
// scft_batch_complete.
int refuse_inflight(const scft_ctx *c)
{
    if (c && c->inflight) {
        set_err("a batched iteration is in flight (call scft_batch_complete first)");
        return SCFT_EINVAL;
    }
    return SCFT_OK;
}

void fill_out(scft_ctx *c, int si, float mf, scft_iter_out *out)
{
    std::memset(out, 0, sizeof(*out));
    const Scalars &s = c->scal_host[si];
    const Slot &sl = c->slots[si];
    out->free_energy = s.F;
    out->rmsdev = s.rms;
    out->dw_rms = s.dw_rms;
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        if (sl.d.ns_sp[sp] <= 0) continue;
        out->qsum[sp] = s.Q[sp];
        const double qmod = std::pow(s.Q[sp], 2);   // SCFT.c:1862-1869
        out->chempot[sp] = (qmod == 0) ? 0.0 : -std::log(qmod) / 2.0;
    }
    out->mf = mf;
}

void free_ctx(scft_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->scal_host) (void)hipHostFree(c->scal_host);
    if (c->rng_host) (void)hipHostFree(c->rng_host);
    for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
    for (hipStream_t st : c->side) (void)hipStreamDestroy(st);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Autotune the contour step for this context's size: time the launches of
// the first step over every chain of a batch-sized set of slots with each
// z candidate and keep the faster; then, for each size class of the active
// set, the fused film step against the split step, and the periodic
// last-axis transform through LDS groups (k_fftx) against the register or
// wave-per-line FFT.
int autotune_z(scft_ctx *c)
{
    const scft_params &p = c->proto;
    const double DS = 1.0 / (p.ns - 1);
    const double DZ = p.lz / c->NZ;
    const double az = DS / (DZ * DZ);
    long long abits;
    std::memcpy(&abits, &az, sizeof(abits));
    scft_derived d0;
    if (scft_derive(&p, &d0) != SCFT_OK) return SCFT_EINVAL;
    const long long v0 = std::max(1LL, (long long)d0.nx * d0.ny * d0.nz);
    const int nt = int(std::min<long long>(c->nslots, std::max(1LL, (kTuneVoxels + v0 - 1) / v0)));
    const int hs = c->NZ - c->NZ / 2;
    const bool dense_ok = hs <= kDenseHMax;
    // candidates: 0 the z transform, 1 the dense z operator (z and x
    // launches apart), 2 the dense z operator fused with x (film step)
    c->zkind = Z_DENSE;
    const bool film_ok = dense_ok && film_fits(c) && film_length(d0.nx) && c->film_mode != 0;
    c->zkind = Z_FFT;
    // SCFT_ZKIND=fft|dense forces the z operator (tests use it to check both);
    // SCFT_FILM=0|1 keeps the dense operator's two launches or fuses them
    const char *env = std::getenv("SCFT_ZKIND");
    const bool force_fft = env && std::strcmp(env, "fft") == 0;
    const bool force_dense = env && std::strcmp(env, "dense") == 0 && dense_ok;
    std::vector<int> cands;
    if (!force_dense) cands.push_back(0);
    if (dense_ok && !force_fft) {
        if (!(film_ok && c->film_mode == 1)) cands.push_back(1);
        if (film_ok) cands.push_back(2);
    }
    // the last-axis form is tuned where k_fftx takes a periodic length of the grid
    // (SCFT_XFFT unset) and the scratch between axes is the slice type
    const bool xtune = c->fftx_mode == 2 && !c->tmp_wide &&
                       ((d0.nx > 1 && fftx_length(d0.nx)) || (c->NY > 1 && fftx_length(c->NY)));
    // result: bits 0-1 the candidate at nt slots, bit 2 + k the film step for
    // size class k, bit 5 + k k_fftx for size class k
    auto apply = [&](int r) {
        const int k = r & 3;
        c->zkind = k == 0 ? Z_FFT : Z_DENSE;
        for (int i = 0; i < 3; ++i) c->film_cls[i] = k != 0 && ((r >> (2 + i)) & 1);
        for (int i = 0; i < 3; ++i) c->xcoal_cls[i] = (r >> (5 + i)) & 1;
    };
    if (cands.size() == 1 && !xtune) {
        apply(cands[0] == 2 ? (2 | (7 << 2)) : cands[0]);
        return SCFT_OK;
    }
    auto key = std::make_tuple(c->NZ, p.mask ? 1 : 0, abits, int(c->tsize), nt, d0.nx,
                               int(cands.size()) * 4 + cands[0], c->NY, xtune ? 1 : 0);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) { apply(it->second); return SCFT_OK; }
    }
    // time per contour step with `ns` slots active (below); SCFT_TUNE_LOG=1:
    // every timing to stderr
    const char *tl = std::getenv("SCFT_TUNE_LOG");
    const bool tune_log = tl && tl[0] == '1';
    auto time_steps = [&](int kind, int ns, double *ms_out, bool xcoal = false) -> int {
        c->zkind = kind == 0 ? Z_FFT : Z_DENSE;
        for (int i = 0; i < 3; ++i) c->film_cls[i] = kind == 2;
        for (int i = 0; i < 3; ++i) c->xcoal_cls[i] = xcoal;
        for (int t = 0; t < c->nslots; ++t) {
            if (t < ns) {
                RETCHK(slot_init(c, c->slots[t], c->proto));
                c->slots[t].active = true;
            } else {
                c->slots[t].active = false;
            }
        }
        RETCHK(build_schedule(c));
        // Contour steps come in two kinds: those where every chain is live
        // (the first n0) and, when some chains are shorter, the rest.  Each
        // kind is timed on a window of up to 20 consecutive steps (the first
        // ones, and those around the middle step) and weighted by its number
        // of steps, so that a form that is slow only on the few steps of
        // short chains is not charged for every step.  Consecutive steps,
        // each reading the slice the one before wrote, as the iteration runs
        // them: one step repeated timed the fused film step 46% slow.
        const int nst = int(c->step_first.size());
        if (nst == 0) {
            *ms_out = 0.0;
            return SCFT_OK;
        }
        auto live = [&](int t) {
            int n = 0;
            for (int si : c->act)
                for (const Chain &ch : c->slots[si].chains) n += ch.nsteps > t;
            return n;
        };
        const int l0 = live(0);
        int n0 = 0;
        for (int t = 0; t < nst; ++t) n0 += live(t) == l0;
        const int tm = nst / 2;
        const bool two = n0 < nst && live(tm) != l0;
        // the launches of steps [t0, t1)
        auto step_launches = [&](int t0, int t1) {
            const int a = c->step_first[t0];
            const int b = t1 < nst ? c->step_first[t1] : int(c->launches.size());
            return std::vector<Launch>(c->launches.begin() + a, c->launches.begin() + b);
        };
        // the window's launches as one graph, as the iteration runs them
        // (eager launches of small kernels would time the host); the fastest
        // of three replays, in us per step
        const bool f32 = c->tsize == 4;
        auto time_graph = [&](int t0, int t1, double *us) -> int {
            std::vector<Launch> zl = step_launches(t0, t1);
            std::swap(zl, c->launches);
            hipGraph_t g = nullptr;
            hipGraphExec_t gx = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                f32 ? launch_steps<float>(c, true) : launch_steps<double>(c, true);
                e = hipStreamEndCapture(c->stream, &g);
            }
            std::swap(zl, c->launches);
            if (e == hipSuccess) e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
            if (e == hipSuccess) e = hipEventCreate(&e0);
            if (e == hipSuccess) e = hipEventCreate(&e1);
            if (e == hipSuccess) e = hipGraphLaunch(gx, c->stream);   // warm
            float ms = 1e30f;
            for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
                float m = 0.f;
                e = hipEventRecord(e0, c->stream);
                if (e == hipSuccess) e = hipGraphLaunch(gx, c->stream);
                if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                if (e == hipSuccess) e = hipEventElapsedTime(&m, e0, e1);
                ms = std::min(ms, m);
            }
            if (e1) (void)hipEventDestroy(e1);
            if (e0) (void)hipEventDestroy(e0);
            if (gx) (void)hipGraphExecDestroy(gx);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                set_err(std::string("autotune timing: ") + hipGetErrorString(e));
                return SCFT_EHIP;
            }
            *us = 1000.0 * ms / (t1 - t0);
            return SCFT_OK;
        };
        double u0 = 0.0, um = 0.0;
        RETCHK(time_graph(0, std::min(n0, 20), &u0));
        if (two) {
            const int a = std::max(n0, tm - 10), b = std::min(nst, a + 20);
            RETCHK(time_graph(a, b, &um));
        }
        const double ms = two ? (n0 * u0 + (nst - n0) * um) / nst : u0;
        if (tune_log)
            std::fprintf(stderr, "scft autotune: NX %d NY %d NZ %d, %d slots, z form %d, k_fftx %d: %.2f us per step "
                         "(%d steps of %.2f, %d of %.2f)\n", d0.nx, c->NY, c->NZ, ns, kind, xcoal ? 1 : 0, ms, n0, u0,
                         two ? nst - n0 : 0, um);
        *ms_out = ms;
        return SCFT_OK;
    };
    // the film step must win by kFilmMargin: where the two forms are within
    // a few percent (a full batch) the choice would follow timing noise
    constexpr double kFilmMargin = 0.97;
    int choice = cands[0];
    double best = 1e30;
    for (int kind : cands) {
        if (cands.size() == 1) break;
        double ms;
        RETCHK(time_steps(kind, nt, &ms));
        if (kind == 2) ms /= kFilmMargin;
        if (ms < best) { best = ms; choice = kind; }
    }
    int result = choice | (choice == 2 ? (4 << 2) : 0);
    if (cands.size() == 1 && choice == 2) result |= 7 << 2;   // the film step forced
    // the film step for the smaller size classes, timed against the split step
    const bool has_film = std::find(cands.begin(), cands.end(), 2) != cands.end();
    const bool has_split = std::find(cands.begin(), cands.end(), 1) != cands.end();
    for (int k = 0; k < 2; ++k) {
        if (choice == 0) break;
        const int ns = std::min(kFilmTuneSlots[k], nt);
        bool f = choice == 2;
        if (ns < nt && has_film && has_split) {
            double t1, t2;
            RETCHK(time_steps(1, ns, &t1));
            RETCHK(time_steps(2, ns, &t2));
            f = t2 < kFilmMargin * t1;
        }
        if (f) result |= 1 << (2 + k);
    }
    // k_fftx against the rule's last-axis form, per size class of the split
    // step (the film step has its own x transform); like the film step it
    // must win by the margin
    if (xtune) {
        std::map<int, bool> at;   // slots timed -> k_fftx chosen
        for (int k = 0; k < 3; ++k) {
            if ((result >> (2 + k)) & 1) continue;
            const int ns = std::min(kFilmTuneSlots[k], nt);
            auto it = at.find(ns);
            if (it == at.end()) {
                const int kind = choice == 0 ? 0 : 1;
                double t1, t2;
                RETCHK(time_steps(kind, ns, &t1, false));
                RETCHK(time_steps(kind, ns, &t2, true));
                it = at.emplace(ns, t2 < kFilmMargin * t1).first;
            }
            if (it->second) result |= 1 << (5 + k);
        }
    }
    for (int t = 0; t < c->nslots; ++t) {
        c->slots[t].active = false;
        c->slots[t].ready = false;
    }
    c->dirty = true;
    apply(result);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = result;
    return SCFT_OK;
}

Slot *get_slot(scft_ctx *c, int slot)
{
    if (!c || slot < 0 || slot >= c->nslots) {
        set_err("bad slot index");
        return nullptr;
    }
    return &c->slots[slot];
}

}  // namespace

// ================================================================== C ABI

extern "C" {

const char *scft_last_error(void) { return g_err.c_str(); }

int scft_batch_create(const scft_params *proto, int nslots, int max_nx, int precision, scft_ctx **out)
{
    if (!proto || !out || nslots < 1 ||
        (precision != SCFT_FP64 && precision != SCFT_FP32_SPLIT && precision != SCFT_FP32_NEAREST)) {
        set_err("scft_batch_create: bad argument");
        return SCFT_EINVAL;
    }
    *out = nullptr;
    scft_ctx *c = new scft_ctx();
    c->precision = precision;
    c->tsize = precision == SCFT_FP64 ? 8 : 4;
    c->proto = *proto;
    int rc = scft_derive(proto, &c->dproto);
    if (rc != SCFT_OK) {
        set_err(rc == SCFT_EVOLUME ? "volumes do not sum to 1" : "invalid parameters");
        delete c;
        return rc;
    }
    c->NY = c->dproto.ny;
    c->NZ = c->dproto.nz;
    c->max_nx = std::max(max_nx, c->dproto.nx);
    const size_t V = size_t(c->max_nx) * c->NY * c->NZ;
    if (V > size_t(1) << 28) {   // kernels address a slice with 32-bit byte offsets
        set_err("grid too large");
        delete c;
        return SCFT_EUNSUPPORTED;
    }
    c->max_V = int(V);
    c->nslots = nslots;
    c->nred = grid_for(V, kRedBlocks);
    const char *g = std::getenv("SCFT_GRAPH");
    c->use_graph = !(g && g[0] == '0');
    // kernel variants (tests force each one): SCFT_XFFT=lds selects the
    // generic LDS FFT for periodic axes, =wave the wave-per-line FFT, =reg
    // the register FFTs at every size, =coal the 32-line LDS groups
    // (k_fftx) wherever the layout allows; SCFT_DENSE=valu the VALU dense z op,
    // SCFT_TMP=fp64 FP64 scratch between axis operators in FP32 mode
    const char *xf = std::getenv("SCFT_XFFT");
    c->fft_wave = !(xf && std::strcmp(xf, "lds") == 0);
    c->fft_reg = !(xf && (std::strcmp(xf, "lds") == 0 || std::strcmp(xf, "wave") == 0));
    c->fft_reg_force = xf && std::strcmp(xf, "reg") == 0;
    c->fftx_mode = !xf ? 2 : (std::strcmp(xf, "coal") == 0 ? 1 : 0);
    const char *pl = std::getenv("SCFT_PLANE");
    c->plane_on = !(pl && pl[0] == '0');
    const char *fm = std::getenv("SCFT_FILM");
    c->film_mode = !fm ? -1 : (fm[0] == '0' ? 0 : 1);
    const char *dm = std::getenv("SCFT_DENSE");
    c->dense_mfma = !(dm && std::strcmp(dm, "valu") == 0);
    c->dense_mfma1 = dm && std::strcmp(dm, "mfma1") == 0;
    const char *tw = std::getenv("SCFT_TMP");
    c->tmp_wide = tw && std::strcmp(tw, "fp64") == 0;
    // SCFT_E2F=1 (FP32-split, grids of at least kE2fMinV points): the MFMA
    // z operator reads E^2 as FP32 inside a block, 4 instead of 8 bytes per
    // voxel-step.  Off by default: at C2 with 192 slots it cut the z
    // launch's bytes from 319 to 241 MB but its time only from 77.5 to
    // 74.6 us (the batched iteration 14.39 -> 14.18 ms), while the per-voxel
    // rounding of E^2 moves F by up to ~1e-6 kT on small grids.
    // SCFT_FP32_SPLIT rounds propagator slices and the scratch between axes
    // to FP32 without bias (store_round), SCFT_FP32_NEAREST to nearest
    c->sround = precision == SCFT_FP32_SPLIT;
    const char *e2 = std::getenv("SCFT_E2F");
    c->e2f = precision != SCFT_FP64 && e2 && e2[0] == '1' && V >= size_t(kE2fMinV);
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            c->ncu = ncu;
    }
    auto fail = [&](int r) {
        free_ctx(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        return fail(SCFT_EHIP);
    }
    const char *px = std::getenv("SCFT_PARX");
    c->par_axes = !(px && px[0] == '0');
    const char *xm = std::getenv("SCFT_XMIX");
    c->xmix = !(xm && xm[0] == '0');
    // SCFT_ZPF=0: the MFMA z operator without the next tile's slice rows in
    // flight (k_zdense_mfma2 instead of k_zdense_mfma2p)
    const char *zp = std::getenv("SCFT_ZPF");
    c->zpf = !(zp && zp[0] == '0');
    // the FFT tiles may use more than 64 KB of LDS on gfx950 (160 KB per workgroup)
    const void *lk[] = {(const void *)k_lineop<double, double>, (const void *)k_lineop<float, double>,
                        (const void *)k_lineop<double, float>, (const void *)k_lineop<float, float>};
    for (const void *f : lk) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    c->slots.resize(nslots);
    if ((rc = ctx_alloc(c, &c->scal, nslots))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void **>(&c->scal_host), sizeof(Scalars) * nslots, hipHostMallocDefault) != hipSuccess) {
        set_err("hipHostMalloc failed");
        return fail(SCFT_ENOMEM);
    }
    std::memset(c->scal_host, 0, sizeof(Scalars) * nslots);
    fill_d(c, reinterpret_cast<double *>(c->scal), 0.0, sizeof(Scalars) / sizeof(double) * nslots);
    for (auto &s : c->slots)
        if ((rc = slot_alloc(c, s))) return fail(rc);
    if ((rc = autotune_z(c))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_err("initialisation failed");
        return fail(SCFT_EHIP);
    }
    *out = c;
    return SCFT_OK;
}

int scft_slot_init(scft_ctx *c, int slot, const scft_params *p)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !p) return SCFT_EINVAL;
    s->active = false;
    s->rng_on = false;   // a new instance starts without a noise stream
    c->dirty = true;
    RETCHK(slot_init(c, *s, *p));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_activate(scft_ctx *c, int slot, int on)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s) return SCFT_EINVAL;
    if (on && !s->ready) {
        set_err("slot not initialised");
        return SCFT_EINVAL;
    }
    if (s->active != (on != 0)) {
        s->active = on != 0;
        c->dirty = true;
    }
    return SCFT_OK;
}

int scft_batch_size(const scft_ctx *c) { return c ? c->nslots : SCFT_EINVAL; }

int scft_slot_bytes(const scft_params *p, int max_nx, int precision, size_t *bytes)
{
    if (!p || !bytes) return SCFT_EINVAL;
    scft_derived d;
    RETCHK(scft_derive(p, &d));
    const size_t V = size_t(std::max(max_nx, d.nx)) * d.ny * d.nz;
    const size_t ts = precision == SCFT_FP64 ? 8 : 4;
    size_t b = 44 * V * sizeof(double);                 // w, p, tot, phitot, mask, E
    if (precision != SCFT_FP64) b += 13 * V * sizeof(float);   // FP32 E^2 (at most)
    size_t nchains = 0;
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        if (d.ns_sp[sp] <= 0) continue;
        const bool homo = sp >= 4 && sp < 14;
        b += (homo ? 1 : 2) * size_t(d.ns_sp[sp]) * V * ts;   // q (and q-dagger)
        nchains += homo ? 1 : 2;
    }
    b += nchains * V * sizeof(double);                  // per-chain axis scratch
    b += (size_t(SCFT_NSPECIES) + 4) * kRedBlocks * sizeof(double);
    *bytes = b;
    return SCFT_OK;
}

int scft_device_memory(size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return SCFT_EINVAL;
    HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
    return SCFT_OK;
}

int scft_create(const scft_params *p, int precision, scft_ctx **out)
{
    RETCHK(scft_batch_create(p, 1, 0, precision, out));
    int rc = scft_slot_init(*out, 0, p);
    if (rc == SCFT_OK) rc = scft_slot_activate(*out, 0, 1);
    if (rc != SCFT_OK) {
        free_ctx(*out);
        *out = nullptr;
    }
    return rc;
}

int scft_destroy(scft_ctx *ctx)
{
    if (!ctx) return SCFT_EINVAL;
    if (ctx->inflight) (void)hipStreamSynchronize(ctx->stream);   // the launch still owns the buffers
    free_ctx(ctx);
    return SCFT_OK;
}

int scft_slot_get_derived(const scft_ctx *c, int slot, scft_derived *out)
{
    if (!c || !out || slot < 0 || slot >= c->nslots) return SCFT_EINVAL;
    *out = c->slots[slot].ready ? c->slots[slot].d : c->dproto;
    return SCFT_OK;
}

int scft_get_derived(const scft_ctx *c, scft_derived *out) { return scft_slot_get_derived(c, 0, out); }

int scft_slot_set_fields(scft_ctx *c, int slot, const double *w, const double *p)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !w || !p || !s->ready) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(s->w, w, sizeof(double) * 13 * s->V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(s->pf, p, sizeof(double) * s->V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_copy_fields(scft_ctx *c, int dst_slot, int src_slot)
{
    RETCHK(refuse_inflight(c));
    Slot *d = get_slot(c, dst_slot);
    Slot *s = get_slot(c, src_slot);
    if (!d || !s || !d->ready || !s->ready || d->V != s->V) {
        set_err("slots differ in size or are not initialised");
        return SCFT_EINVAL;
    }
    HIPCHK(hipMemcpyAsync(d->w, s->w, sizeof(double) * 13 * s->V, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d->pf, s->pf, sizeof(double) * s->V, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_get_fields(scft_ctx *c, int slot, double *w, double *p)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !w || !p || !s->ready) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(w, s->w, sizeof(double) * 13 * s->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(p, s->pf, sizeof(double) * s->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_set_mask(scft_ctx *c, int slot, const double *mask)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !mask || !s->ready) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(s->mask, mask, sizeof(double) * 3 * s->V, hipMemcpyHostToDevice, c->stream));
    // brush q starts from the mask every iteration (SCFT.c:2441); slices
    // s >= 1 never overwrite slice 0, so setting it once is enough
    for (int b = 0; b < 3; ++b) {
        if (s->d.ns_sp[14 + b] <= 0) continue;
        if (c->tsize == 8)
            hipLaunchKernelGGL(k_convert<double>, dim3(grid_for(s->V)), dim3(kThreads), 0, c->stream,
                               s->mask + (size_t)b * s->V, (double *)s->q[14 + b], s->V);
        else
            hipLaunchKernelGGL(k_convert<float>, dim3(grid_for(s->V)), dim3(kThreads), 0, c->stream,
                               s->mask + (size_t)b * s->V, (float *)s->q[14 + b], s->V);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_get_densities(scft_ctx *c, int slot, double *phi, double *phi_tot)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !phi || !s->ready) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(phi, s->tot, sizeof(double) * 13 * s->V, hipMemcpyDeviceToHost, c->stream));
    if (phi_tot)
        HIPCHK(hipMemcpyAsync(phi_tot, s->phitot, sizeof(double) * s->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_slot_get_propagator(scft_ctx *c, int slot, int sp, int which, int idx, double *out)
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !out || !s->ready) return SCFT_EINVAL;
    if (sp < 0 || sp >= SCFT_NSPECIES || s->d.ns_sp[sp] <= 0 || idx < 0 || idx >= s->d.ns_sp[sp]) {
        set_err("scft_get_propagator: no such slice");
        return SCFT_EINVAL;
    }
    const void *base = which ? s->qd[sp] : s->q[sp];
    const Chain *ch = find_chain(*s, base);
    if (!ch) {
        set_err("scft_get_propagator: no chain for the slice");
        return SCFT_EINVAL;
    }
    const double *fac = slice_factor(*s, *ch, idx);
    const char *src = static_cast<const char *>(base) + (size_t)idx * s->V * c->tsize;
    double *tmp = nullptr;
    RETCHK(dev_alloc(&tmp, s->V));
    if (c->tsize == 8)
        hipLaunchKernelGGL(k_widen<double>, dim3(grid_for(s->V)), dim3(kThreads), 0, c->stream, (const double *)src, fac, tmp, s->V);
    else
        hipLaunchKernelGGL(k_widen<float>, dim3(grid_for(s->V)), dim3(kThreads), 0, c->stream, (const float *)src, fac, tmp, s->V);
    hipError_t e = hipMemcpyAsync(out, tmp, sizeof(double) * s->V, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) {
        set_err(std::string("get_propagator: ") + hipGetErrorString(e));
        return SCFT_EHIP;
    }
    return SCFT_OK;
}

int scft_set_fields(scft_ctx *c, const double *w, const double *p) { return scft_slot_set_fields(c, 0, w, p); }
int scft_get_fields(scft_ctx *c, double *w, double *p) { return scft_slot_get_fields(c, 0, w, p); }
int scft_set_mask(scft_ctx *c, const double *m) { return scft_slot_set_mask(c, 0, m); }
int scft_get_densities(scft_ctx *c, double *phi, double *t) { return scft_slot_get_densities(c, 0, phi, t); }
int scft_get_propagator(scft_ctx *c, int sp, int which, int s, double *out)
{
    return scft_slot_get_propagator(c, 0, sp, which, s, out);
}

int scft_propagate(scft_ctx *c, unsigned long iter)
{
    RETCHK(refuse_inflight(c));
    if (!c) return SCFT_EINVAL;
    if (c->dirty) RETCHK(build_schedule(c));
    for (int si : c->act) c->slots[si].iter = iter;
    RETCHK(set_iteration_args(c, 0.0f, nullptr));
    return enqueue(c, 1);
}

int scft_densities(scft_ctx *c, double *qsum, double *chempot)
{
    RETCHK(refuse_inflight(c));
    if (!c) return SCFT_EINVAL;
    if (c->dirty) RETCHK(build_schedule(c));
    RETCHK(enqueue(c, 2 | 8));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->act.empty()) {
        scft_iter_out o;
        fill_out(c, c->act[0], 0.0f, &o);
        for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
            if (qsum) qsum[sp] = o.qsum[sp];
            if (chempot) chempot[sp] = o.chempot[sp];
        }
    }
    return SCFT_OK;
}

static int check_rng(scft_ctx *c)
{
    for (int si : c->act) {
        const Slot &s = c->slots[si];
        if (s.rng_on && s.rng_need && c->rng_host[si].short_of_pairs) {
            set_err("device noise stream of slot " + std::to_string(si) + ": " +
                    std::to_string(c->rng_host[si].found) + " accepted pairs among the candidates, " +
                    std::to_string(s.rng_need) + " needed");
            return SCFT_ERNG;
        }
    }
    return SCFT_OK;
}

static int upload_noise(scft_ctx *c, float mf, const double *noise, const double **dev)
{
    *dev = nullptr;
    if (c->slots[0].rng_on) {   // the device draws the deviates
        if (noise) {
            set_err("noise must be NULL while the device noise stream is on");
            return SCFT_EINVAL;
        }
        if (c->act.size() != 1 || c->act[0] != 0) {
            set_err("the device noise stream needs slot 0 active alone");
            return SCFT_EINVAL;
        }
        if (mf != 0.0f) *dev = c->slots[0].noise;
        return SCFT_OK;
    }
    if (mf == 0.0f) return SCFT_OK;
    if (!noise) {
        set_err("noise is required when mf != 0");
        return SCFT_EINVAL;
    }
    if (c->act.size() != 1) {
        set_err("noise is supported for a single active slot");
        return SCFT_EINVAL;
    }
    Slot &s = c->slots[c->act[0]];
    const size_t n = size_t(s.d.n_active) * s.V;
    if (!s.noise) RETCHK(ctx_alloc(c, &s.noise, size_t(13) * c->max_V));
    HIPCHK(hipMemcpyAsync(s.noise, noise, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    *dev = s.noise;
    return SCFT_OK;
}

int scft_update_fields(scft_ctx *c, unsigned long iter, float mf, const double *noise, scft_iter_out *out)
{
    RETCHK(refuse_inflight(c));
    (void)iter;
    if (!c) return SCFT_EINVAL;
    if (c->dirty) RETCHK(build_schedule(c));
    const double *nd = nullptr;
    RETCHK(upload_noise(c, mf, noise, &nd));
    RETCHK(set_iteration_args(c, mf, nd));
    RETCHK(enqueue(c, 4 | 8));
    HIPCHK(hipStreamSynchronize(c->stream));
    RETCHK(check_rng(c));
    if (out && !c->act.empty()) fill_out(c, c->act[0], mf, out);
    return SCFT_OK;
}

int scft_iterate(scft_ctx *c, unsigned long iter, float mf, const double *noise, scft_iter_out *out)
{
    RETCHK(refuse_inflight(c));
    if (!c) return SCFT_EINVAL;
    if (c->dirty) RETCHK(build_schedule(c));
    for (int si : c->act) c->slots[si].iter = iter;
    const double *nd = nullptr;
    RETCHK(upload_noise(c, mf, noise, &nd));
    RETCHK(run_iteration(c, mf, nd));
    RETCHK(check_rng(c));
    if (out && !c->act.empty()) fill_out(c, c->act[0], mf, out);
    return SCFT_OK;
}

// Seed the device replay of a slot's noise stream with a drand48 state
// (seed48 layout: xsubi[0] the low 16 bits).  Sizes the candidate pairs
// per iteration: need / (pi/4) plus `SCFT_RNG_MARGIN` (default 8) times
// sqrt(need) + 256 spare pairs, about 17 standard deviations of the
// accepted count; fewer accepted pairs than needed is reported as
// SCFT_ERNG and leaves the state unadvanced.  Re-initialising the slot
// (scft_slot_init) turns its stream off.
int scft_slot_rng_seed(scft_ctx *c, int slot, const unsigned short xsubi[3])
{
    RETCHK(refuse_inflight(c));
    Slot *sp = get_slot(c, slot);
    if (!sp || !xsubi || !sp->ready) return SCFT_EINVAL;
    Slot &s = *sp;
    const unsigned long long need = (unsigned long long)s.d.n_active * (unsigned long long)s.V;
    double k = 8.0;
    if (const char *m = std::getenv("SCFT_RNG_MARGIN")) k = std::atof(m);
    const double pairs = (double(need) + k * std::sqrt(double(need)) + (k > 0 ? 256.0 : 0.0)) / (M_PI / 4.0);
    const unsigned long long nchunk = need ? (unsigned long long)std::ceil(std::max(pairs, 1.0) / kRngChunk) : 0;
    if (need >= (1ull << 31) || nchunk >= (1ull << 30)) {
        set_err("noise stream too long for one iteration");
        return SCFT_EUNSUPPORTED;
    }
    if (!c->rng) {
        RETCHK(ctx_alloc(c, &c->rng, size_t(c->nslots)));
        RETCHK(ctx_alloc(c, &c->rng_jump, 1));
        if (hipHostMalloc(reinterpret_cast<void **>(&c->rng_host), sizeof(RngState) * c->nslots,
                          hipHostMallocDefault) != hipSuccess) {
            c->rng_host = nullptr;
            set_err("hipHostMalloc failed");
            return SCFT_ENOMEM;
        }
        std::memset(c->rng_host, 0, sizeof(RngState) * c->nslots);
        RngJump J;
        unsigned long long a = kLcgA, cc = kLcgC;   // x -> a x + cc advances 2^b draws
        for (int b = 0; b < 48; ++b) {
            J.a[b] = a;
            J.c[b] = cc;
            cc = (a * cc + cc) & kLcgMask;
            a = (a * a) & kLcgMask;
        }
        HIPCHK(hipMemcpyAsync(c->rng_jump, &J, sizeof(J), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (int(nchunk) > s.rng_cap) {
        RETCHK(ctx_alloc(c, &s.rng_counts, size_t(nchunk)));
        RETCHK(ctx_alloc(c, &s.rng_offsets, size_t(nchunk)));
        s.rng_cap = int(nchunk);
    }
    if (!s.noise) RETCHK(ctx_alloc(c, &s.noise, size_t(13) * c->max_V));
    RngState st{};
    st.x = ((unsigned long long)xsubi[2] << 32) | ((unsigned long long)xsubi[1] << 16) | xsubi[0];
    st.xn = st.x;
    HIPCHK(hipMemcpyAsync(c->rng + slot, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->rng_host[slot] = st;
    s.rng_nchunk = int(nchunk);
    s.rng_need = need;
    s.rng_on = true;
    c->dirty = true;   // the iteration now includes the slot's generator
    return SCFT_OK;
}

// The slot's drand48 state after its last completed iteration.
int scft_slot_rng_state(scft_ctx *c, int slot, unsigned short xsubi[3])
{
    RETCHK(refuse_inflight(c));
    Slot *s = get_slot(c, slot);
    if (!s || !xsubi) return SCFT_EINVAL;
    if (!s->rng_on) {
        set_err("the slot's device noise stream is off (scft_slot_rng_seed)");
        return SCFT_EINVAL;
    }
    RngState st;
    HIPCHK(hipMemcpyAsync(&st, c->rng + slot, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    xsubi[0] = (unsigned short)(st.x & 0xffff);
    xsubi[1] = (unsigned short)((st.x >> 16) & 0xffff);
    xsubi[2] = (unsigned short)((st.x >> 32) & 0xffff);
    return SCFT_OK;
}

int scft_rng_seed(scft_ctx *c, const unsigned short xsubi[3])
{
    return scft_slot_rng_seed(c, 0, xsubi);
}

int scft_rng_state(scft_ctx *c, unsigned short xsubi[3])
{
    return scft_slot_rng_state(c, 0, xsubi);
}

int scft_batch_launch(scft_ctx *c, const unsigned long *iters)
{
    if (!c) return SCFT_EINVAL;
    RETCHK(refuse_inflight(c));
    if (c->dirty) RETCHK(build_schedule(c));
    for (int si : c->act) {
        Slot &s = c->slots[si];
        s.iter = iters ? iters[si] : 1ul << 40;
        if (s.iter <= s.p.relaxation && !s.rng_on) {   // MF = 1 needs the slot's stream (SCFT.c:1515-1522)
            set_err("slot " + std::to_string(si) + " is within RELAXATION but has no noise stream (scft_slot_rng_seed)");
            return SCFT_EINVAL;
        }
    }
    RETCHK(run_iteration(c, -1.0f, nullptr, false));
    c->inflight = true;
    return SCFT_OK;
}

int scft_batch_complete(scft_ctx *c, scft_iter_out *outs)
{
    if (!c) return SCFT_EINVAL;
    if (!c->inflight) {
        set_err("no batched iteration in flight (scft_batch_launch)");
        return SCFT_EINVAL;
    }
    c->inflight = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    RETCHK(check_rng(c));
    if (outs)
        for (int si : c->act) {
            const Slot &s = c->slots[si];
            fill_out(c, si, (s.rng_on && s.iter <= s.p.relaxation) ? 1.0f : 0.0f, &outs[si]);
        }
    return SCFT_OK;
}

int scft_batch_iterate(scft_ctx *c, const unsigned long *iters, scft_iter_out *outs)
{
    RETCHK(scft_batch_launch(c, iters));
    return scft_batch_complete(c, outs);
}

int scft_batch_query(scft_ctx *c)
{
    if (!c) return SCFT_EINVAL;
    if (!c->inflight) {
        set_err("no batched iteration in flight (scft_batch_launch)");
        return SCFT_EINVAL;
    }
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_err(std::string("hipStreamQuery: ") + hipGetErrorString(e));
    return SCFT_EHIP;
}

int scft_sync(scft_ctx *c)
{
    if (!c) return SCFT_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

// Time every launch of one kernel kind of the current schedule (one
// iteration's worth), `reps` times, between two events on the context's
// stream; report the mean launch duration and the algorithmic bytes and
// flops of the mean launch.  kind: 0 = z operator, 1 = last (x) axis.
int scft_time_launches(scft_ctx *c, int kind, int reps, double *avg_us, double *bytes, double *flops,
                       double *voxel_steps, int *nlaunch)
{
    if (!c || reps < 1 || !avg_us || !bytes || !flops || !voxel_steps || !nlaunch) return SCFT_EINVAL;
    RETCHK(refuse_inflight(c));
    if (c->dirty) RETCHK(build_schedule(c));
    std::vector<Launch> sel;
    double B = 0.0, Fl = 0.0, U = 0.0;
    const double ts = double(c->tsize);
    for (const Launch &ln : c->launches) {
        const bool is_z = (ln.kind == LK_DENSE || ln.kind == LK_DENSE_MFMA || ln.kind == LK_DENSE_MFMA2) ||
                          ln.kind == LK_PLANE || ln.kind == LK_FILM ||
                          ((ln.kind == LK_LINEOP || ln.kind == LK_FFTWAVE || ln.kind == LK_FFTREG ||
                            ln.kind == LK_FFTREG2 || ln.kind == LK_FFTX || ln.kind == LK_FFTMIX) && ln.pos == 0);
        if ((kind == 0) != is_z) continue;
        if (kind == 1 && ln.pos != 2 && ln.pos != 3) continue;
        sel.push_back(ln);
        // algorithmic traffic: read src and the Boltzmann factors (first axis:
        // one FP64 array, two at a block boundary), write dst; slices are T,
        // the scratch between axes T (FP64 when wide)
        const double in_b = (ln.pos == 0 || ln.pos == 3 || !c->tmp_wide) ? ts : 8.0;
        const double out_b = (ln.pos == 2 || ln.pos == 3 || !c->tmp_wide) ? ts : 8.0;
        auto e_bytes = [](const double *e1, const double *e2, const double *ep, const float *e2f = nullptr) {
            if (e2f) return 4.0 + (ep ? 8.0 : 0.0);
            return (e1 ? (e2 && e2 != e1 ? 16.0 : 8.0) : 0.0) + (ep ? 8.0 : 0.0);
        };
        if (ln.kind == LK_FILM) {
            std::vector<FilmOp> ops(ln.count);
            HIPCHK(hipMemcpy(ops.data(), c->filmops + ln.first, ln.count * sizeof(FilmOp), hipMemcpyDeviceToHost));
            for (auto &op : ops) {
                const double elems = double(ln.L) * kFilmN;
                U += elems;
                B += elems * (in_b + out_b + e_bytes(op.epre, op.epre2, nullptr, op.e2f));
                // z: NZ^2/2 FMA per line of NZ; x: 2 x 5 L log2 L per packed line pair, NZ/2 pairs
                Fl += 2.0 * ln.L * (2.0 * (kFilmN / 2.0) * (kFilmN / 2.0)) +
                      (kFilmN / 2.0) * 2.0 * 5.0 * ln.L * std::log2(double(ln.L));
            }
        } else if (ln.kind == LK_PLANE) {
            std::vector<PlaneOp> ops(ln.count);
            HIPCHK(hipMemcpy(ops.data(), c->planeops + ln.first, ln.count * sizeof(PlaneOp), hipMemcpyDeviceToHost));
            for (auto &op : ops) {
                const double elems = double(kPlaneN) * kPlaneN;
                U += elems;
                B += elems * (in_b + out_b + e_bytes(op.epre, op.epre2, nullptr));
                // both axes: 2 x 5 L log2 L per packed line pair, 32 pairs per axis
                Fl += 2.0 * (kPlaneN / 2.0) * 2.0 * 5.0 * kPlaneN * std::log2(double(kPlaneN));
            }
        } else if (ln.kind == LK_DENSE || ln.kind == LK_DENSE_MFMA || ln.kind == LK_DENSE_MFMA2) {
            std::vector<DenseOp> ops(ln.count);
            HIPCHK(hipMemcpy(ops.data(), c->denseops + ln.first, ln.count * sizeof(DenseOp), hipMemcpyDeviceToHost));
            for (auto &op : ops) {
                const double elems = double(op.nrows) * op.n;
                U += elems;
                B += elems * (in_b + out_b +
                              e_bytes(op.epre, op.epre2, op.epost, ln.kind == LK_DENSE_MFMA2 ? op.e2f : nullptr));
                Fl += 2.0 * op.nrows * (double(op.hs) * op.hs + double(op.ha) * op.ha);
            }
        } else {
            std::vector<LineOp> ops(ln.count);
            HIPCHK(hipMemcpy(ops.data(), c->lineops + ln.first, ln.count * sizeof(LineOp), hipMemcpyDeviceToHost));
            for (auto &op : ops) {
                const double elems = double(op.nlines) * op.n;
                U += elems;
                B += elems * (in_b + out_b + e_bytes(op.epre, op.epre2, op.epost));
                // radix-2-equivalent count of a complex transform pair: 2 x 5 L log2 L per packed line pair
                Fl += (op.nlines / 2.0) * 2.0 * 5.0 * op.L * std::log2(double(op.L));
            }
        }
    }
    *nlaunch = int(sel.size());
    if (sel.empty()) {
        *avg_us = *bytes = *flops = *voxel_steps = 0.0;
        return SCFT_OK;
    }
    std::swap(sel, c->launches);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool f32 = c->tsize == 4;
    f32 ? launch_steps<float>(c) : launch_steps<double>(c);   // warm
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) f32 ? launch_steps<float>(c) : launch_steps<double>(c);
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    std::swap(sel, c->launches);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double nl = double(*nlaunch);
    *avg_us = 1000.0 * ms / (reps * nl);
    *bytes = B / nl;
    *flops = Fl / nl;
    *voxel_steps = U / nl;
    return SCFT_OK;
}

int scft_get_schedule(const scft_ctx *c, char *buf, int len)
{
    if (!c || !buf || len <= 0) return SCFT_EINVAL;
    int n[kLaunchKinds] = {}, npf = 0;
    for (auto &ln : c->launches) {
        if (ln.kind >= 0 && ln.kind < kLaunchKinds) n[ln.kind]++;
        npf += ln.kind == LK_DENSE_MFMA2 && ln.pf;
    }
    std::snprintf(buf, len,
                  "z=%s launches/iter: lineop=%d dense=%d fftwave=%d fftreg=%d dense_mfma=%d dense_mfma2=%d graph=%d "
                  "fftreg2=%d plane=%d film=%d fftx=%d mfma2_pf=%d fftmix=%d",
                  c->zkind == Z_DENSE ? "dense" : "fft", n[LK_LINEOP], n[LK_DENSE], n[LK_FFTWAVE], n[LK_FFTREG],
                  n[LK_DENSE_MFMA], n[LK_DENSE_MFMA2], c->use_graph ? 1 : 0, n[LK_FFTREG2], n[LK_PLANE],
                  n[LK_FILM], n[LK_FFTX], npf, n[LK_FFTMIX]);
    return SCFT_OK;
}

}  // extern "C"

// ============================================== one contour step, no allocation
//
// scft_step_noalloc: one contour step of one chain on caller device arrays,
// for callers that cannot allocate device memory.  Its tables, its Boltzmann
// factor, its transform-op blocks and the scratch between axis operators
// live in `scratch`, a file-scope device array that every GPU running the
// library holds until the process ends.  Host code never uses the array's
// name as an address: the per-device address comes from
// hipGetSymbolAddress, resolved once per device with that device current.

namespace {

constexpr size_t kNaMaxV = size_t(1) << 18;          // voxels per call
constexpr int kNaMaxL = 1024;                         // transform length per axis
constexpr size_t kNaOpsBytes = 1024;                  // 3 LineOp blocks
constexpr size_t kNaTableBytes = size_t(3) * kNaMaxL * (sizeof(double) + sizeof(cpx<double>));
constexpr size_t kNaScratchBytes = kNaOpsBytes + kNaTableBytes + 2 * kNaMaxV * sizeof(double);
static_assert(3 * sizeof(LineOp) <= kNaOpsBytes, "op blocks exceed their region");
static_assert(kNaOpsBytes % 256 == 0 && kNaTableBytes % 256 == 0, "regions stay 256-byte aligned");

std::mutex g_na_mu;
void *g_na_ptr[64] = {};                              // scratch address per device
bool g_na_attr = false;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace scft {

// K and the twiddles of one axis (the tables get_fft computes on the host).
__global__ void __launch_bounds__(kThreads) k_na_tables(double *K, cpx<double> *tw, int L, double alpha)
{
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < L; m += gridDim.x * blockDim.x) {
        const double km = 2.0 * M_PI * (L / 2.0 - fabs(m - L / 2.0)) / L;
        K[m] = exp(-alpha * km * km);
        const double a = 2.0 * M_PI * double(m) / double(L);
        tw[m] = {cos(a), -sin(a)};
    }
}

// E = exp(-DS w / 2) of the chain's field (k_prep's formula).
__global__ void __launch_bounds__(kThreads) k_na_factor(const double *w, double *E, double ds, int V)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x)
        E[i] = exp(ds * (-w[i] / 2.0));
}

// Writes a transform-op block (passed by value) into device memory.
__global__ void k_na_put_op(LineOp op, LineOp *dst) { *dst = op; }

}  // namespace scft

namespace {

int na_scratch(unsigned char **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_err("device index out of range");
        return SCFT_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_attr) {
        HIPCHK(hipFuncSetAttribute((const void *)k_lineop<double, double>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        g_na_attr = true;
    }
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return SCFT_EHIP;
        }
        g_na_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_na_ptr[dev]);
    return SCFT_OK;
}

}  // namespace

extern "C" {

int scft_step_noalloc(const scft_params *p, const double *w_dev, const double *q_in_dev, double *q_out_dev,
                      void *stream)
{
    if (!p || !w_dev || !q_in_dev || !q_out_dev) return SCFT_EINVAL;
    scft_derived d;
    RETCHK(scft_derive(p, &d));
    const size_t V = size_t(d.nx) * d.ny * d.nz;
    if (V > kNaMaxV) {
        set_err("scft_step_noalloc: " + std::to_string(V) + " voxels exceed the limit of " + std::to_string(kNaMaxV));
        return SCFT_EUNSUPPORTED;
    }
    const int dir = p->mask ? 1 : 0;
    const double DS = 1.0 / (p->ns - 1);
    const double DX = p->lx / d.nx, DY = p->ly / d.ny, DZ = p->lz / d.nz;
    struct Axis { int axis, L; double alpha; Radices rad; };
    std::vector<Axis> axes;
    axes.push_back({2, dir ? 2 * (d.nz + 1) : d.nz, DS / (DZ * DZ), {}});
    if (d.ny > 1) axes.push_back({1, d.ny, DS / (DY * DY), {}});
    if (d.nx > 1) axes.push_back({0, d.nx, DS / (DX * DX), {}});
    for (auto &a : axes) {
        if (a.L > kNaMaxL || !plan_radices(a.L, a.rad)) {
            set_err("scft_step_noalloc: transform length " + std::to_string(a.L) + " is not supported");
            return SCFT_EUNSUPPORTED;
        }
    }
    unsigned char *base = nullptr;
    RETCHK(na_scratch(&base));
    hipStream_t st = static_cast<hipStream_t>(stream);
    LineOp *ops = reinterpret_cast<LineOp *>(base);
    unsigned char *tables = base + kNaOpsBytes;
    double *E = reinterpret_cast<double *>(base + kNaOpsBytes + kNaTableBytes);
    double *tmp = E + kNaMaxV;
    hipLaunchKernelGGL(k_na_factor, dim3(grid_for(V, 512)), dim3(kThreads), 0, st, w_dev, E, DS, int(V));
    const int nax = int(axes.size());
    for (int ai = 0; ai < nax; ++ai) {
        const Axis &a = axes[ai];
        double *K = reinterpret_cast<double *>(tables + size_t(ai) * kNaMaxL * (sizeof(double) + sizeof(cpx<double>)));
        cpx<double> *tw = reinterpret_cast<cpx<double> *>(K + kNaMaxL);
        hipLaunchKernelGGL(k_na_tables, dim3(grid_for(a.L)), dim3(kThreads), 0, st, K, tw, a.L, a.alpha);
        LineOp op{};
        op.src = ai == 0 ? static_cast<const void *>(q_in_dev) : static_cast<const void *>(tmp);
        op.dst = ai == nax - 1 ? static_cast<void *>(q_out_dev) : static_cast<void *>(tmp);
        op.epre = ai == 0 ? E : nullptr;
        op.epre2 = nullptr;
        op.epost = ai == nax - 1 ? E : nullptr;
        op.K = K;
        op.tw = tw;
        op.L = a.L;
        op.npass = a.rad.npass;
        for (int i = 0; i < kMaxPass; ++i) op.radix[i] = i < a.rad.npass ? a.rad.r[i] : 1;
        if (a.axis == 2) {
            op.n = d.nz;
            op.dirichlet = dir;
            op.nlines = d.nx * d.ny;
            op.c1 = 1; op.s0 = 0; op.s1 = d.nz; op.stride = 1;
        } else if (a.axis == 1) {
            op.n = d.ny;
            op.nlines = d.nx * d.nz;
            op.c1 = d.nz; op.s0 = 1; op.s1 = d.nz * d.ny; op.stride = d.nz;
        } else {
            op.n = d.nx;
            op.nlines = d.ny * d.nz;
            op.c1 = d.ny * d.nz; op.s0 = 1; op.s1 = 0; op.stride = d.ny * d.nz;
        }
        op.pairs = (op.s0 == 1 && op.c1 % 2 == 0 && op.s1 % 2 == 0 && op.stride % 2 == 0) ? 1 : 0;
        hipLaunchKernelGGL(k_na_put_op, dim3(1), dim3(1), 0, st, op, ops + ai);
        const int P = lines_P(a.L);
        const size_t lds = size_t(2) * P * a.L * sizeof(cpx<double>);
        hipLaunchKernelGGL((k_lineop<double, double>), dim3((op.nlines + 2 * P - 1) / (2 * P), 1), dim3(kThreads), lds,
                           st, static_cast<const LineOp *>(ops + ai), P);
    }
    HIPCHK(hipGetLastError());
    return SCFT_OK;
}

}  // extern "C"
