// scft_lib.hip — libscft: the SCFT.c iteration loop on one MI355X (gfx950).
//
// Host side of the C ABI declared in include/scft.h.  One context = one SCFT
// instance with device-resident fields, propagators (stored [s][idx], one
// contiguous slice per contour point) and densities, driven on one stream.
// The iteration is a fixed launch sequence:
//   k_prep      Boltzmann factors E_f = exp(-DS w_f / 2) (+ dbk1 seed start)
//   k_lineop    per contour step and axis: every propagator chain of the
//               iteration advances one step in the same launch
//   k_qsum / k_qfinal       Q_p = mean q_p(NS_p - 1)            (FP64)
//   k_density   Simpson contour sums, volume scaling, totals, the free-
//               energy integrand and (phi_tot - 1)^2 per voxel
//   k_update    funUpdate per voxel                               (FP64)
//   k_final / k_pshift      F, RMSDev, mean p, ||dw||; p -= mean(p)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
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
std::vector<double> axis_k(int L, double alpha)
{
    std::vector<double> k(L);
    for (int m = 0; m < L; ++m) {
        const double km = 2.0 * M_PI * (L / 2.0 - std::fabs(m - L / 2.0)) / L;
        k[m] = std::exp(-alpha * std::pow(km, 2));
    }
    return k;
}

template <typename T>
std::vector<cpx<T>> twiddles(int L)
{
    std::vector<cpx<T>> t(L);
    for (int e = 0; e < L; ++e) {
        const double a = 2.0 * M_PI * double(e) / double(L);
        t[e] = {T(std::cos(a)), T(-std::sin(a))};
    }
    return t;
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

struct Axis {
    int n, L, dirichlet, nlines, c1, s0, s1, stride;
    Radices rad;
    void *K = nullptr;   // T[L]
    void *tw = nullptr;  // cpx<T>[L]
};

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual int propagate(unsigned long iter) = 0;
    virtual int densities() = 0;
    virtual int update(float mf, const double *noise_dev) = 0;
    virtual int set_mask(const double *mask_dev) = 0;
    virtual int get_propagator(int sp, int which, int s, double *host) = 0;
};

}  // namespace

struct scft_ctx {
    scft_params p{};
    scft_derived d{};
    int precision = SCFT_FP64;
    int NX = 0, NY = 0, NZ = 0, V = 0;
    int nred = 1;                 // blocks of each reduction kernel
    hipStream_t stream = nullptr;
    double *w = nullptr, *pf = nullptr, *tot = nullptr, *phitot = nullptr, *mask = nullptr;
    double *qpart = nullptr, *dpart = nullptr, *upart = nullptr;
    double *noise = nullptr;
    Scalars *scal = nullptr;
    Scalars *scal_host = nullptr; // pinned
    FinArgs *fin = nullptr;
    UpdArgs upd_host{};
    UpdArgs *upd = nullptr;
    std::unique_ptr<EngineBase> eng;
};

namespace {

template <typename T>
class Engine final : public EngineBase {
public:
    explicit Engine(scft_ctx *c) : c_(c) {}
    ~Engine() override
    {
        for (void *ptr : allocs_) (void)hipFree(ptr);
    }

    int init()
    {
        scft_ctx &c = *c_;
        const scft_derived &d = c.d;
        const size_t V = c.V;
        // ---- propagator storage and chains
        RETCHK(alloc(&E_, 13 * V));
        for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
            const int ns = d.ns_sp[sp];
            if (ns <= 0) continue;
            RETCHK(alloc(&q_[sp], (size_t)ns * V));
            fill(q_[sp], T(1), V);                       // q[0] = 1 (SCFT.c:881-981)
            const bool homo = sp >= 4 && sp < 14;
            if (!homo) {
                RETCHK(alloc(&qd_[sp], (size_t)ns * V));
                fill(qd_[sp], T(1), V);
            } else {
                qd_[sp] = q_[sp];                        // q-dagger = q for homopolymers
            }
        }
        // chains in the reference's species order (SCFT.c:1092-1456)
        for (int h = 0; h < 10; ++h)
            if (d.ns_sp[4 + h] > 0) add_chain(q_[4 + h], d.ns_sp[4 + h] - 1, h, h, 1 << 30);
        for (int b = 0; b < 4; ++b) {
            if (d.ns_sp[b] <= 0) continue;
            const int fa = kDbkField[b][0], fb = kDbkField[b][1];
            add_chain(q_[b], d.ns_sp[b] - 1, fa, fb, d.ns_x[b] - 1);   // A then B (1250-1256)
            add_chain(qd_[b], d.ns_sp[b] - 1, fb, fa, d.ns_y[b]);      // B then A (1263-1269)
        }
        for (int s = 0; s < 3; ++s) {
            const int sp = 14 + s;
            if (d.ns_sp[sp] <= 0) continue;
            add_chain(q_[sp], d.ns_sp[sp] - 1, 10 + s, 10 + s, 1 << 30);  // mask start (2441)
            add_chain(qd_[sp], d.ns_sp[sp] - 1, 10 + s, 10 + s, 1 << 30);
        }
        for (auto &ch : chains_) RETCHK(alloc(&ch.tmp, V));

        // ---- axes
        const double DS = 1.0 / (c.p.ns - 1);
        const double DX = c.p.lx / c.NX, DY = c.p.ly / c.NY, DZ = c.p.lz / c.NZ;
        axes_.clear();
        {   // z (contiguous lines), always applied
            Axis a{};
            a.n = c.NZ;
            a.dirichlet = c.p.mask ? 1 : 0;
            a.L = a.dirichlet ? 2 * (c.NZ + 1) : c.NZ;
            a.nlines = c.NX * c.NY;
            a.c1 = 1; a.s0 = 0; a.s1 = c.NZ; a.stride = 1;
            RETCHK(setup_axis(a, DS / (DZ * DZ)));
            axes_.push_back(a);
        }
        if (c.NY > 1) {
            Axis a{};
            a.n = a.L = c.NY;
            a.nlines = c.NX * c.NZ;
            a.c1 = c.NZ; a.s0 = 1; a.s1 = c.NZ * c.NY; a.stride = c.NZ;
            RETCHK(setup_axis(a, DS / (DY * DY)));
            axes_.push_back(a);
        }
        if (c.NX > 1) {
            Axis a{};
            a.n = a.L = c.NX;
            a.nlines = c.NY * c.NZ;
            a.c1 = c.NY * c.NZ; a.s0 = 1; a.s1 = 0; a.stride = c.NY * c.NZ;
            RETCHK(setup_axis(a, DS / (DX * DX)));
            axes_.push_back(a);
        }

        // ---- step schedule: for each step t, for each axis, one op per live chain
        maxsteps_ = 0;
        for (auto &ch : chains_) maxsteps_ = std::max(maxsteps_, ch.nsteps);
        std::vector<LineOp<T>> ops;
        launches_.clear();
        const int nax = int(axes_.size());
        for (int t = 0; t < maxsteps_; ++t) {
            for (int ai = 0; ai < nax; ++ai) {
                const Axis &ax = axes_[ai];
                Launch ln{};
                ln.first = int(ops.size());
                int maxl = 0;
                for (auto &ch : chains_) {
                    if (ch.nsteps <= t) continue;
                    const int f = (t < ch.switch_at) ? ch.fa : ch.fb;
                    LineOp<T> op{};
                    op.src = (ai == 0) ? ch.base + (size_t)t * V : ch.tmp;
                    op.dst = (ai == nax - 1) ? ch.base + (size_t)(t + 1) * V : ch.tmp;
                    op.epre = (ai == 0) ? E_ + (size_t)f * V : nullptr;
                    op.epost = (ai == nax - 1) ? E_ + (size_t)f * V : nullptr;
                    op.K = static_cast<const T *>(ax.K);
                    op.tw = static_cast<const cpx<T> *>(ax.tw);
                    op.n = ax.n; op.L = ax.L; op.dirichlet = ax.dirichlet;
                    op.nlines = ax.nlines;
                    op.c1 = ax.c1; op.s0 = ax.s0; op.s1 = ax.s1; op.stride = ax.stride;
                    op.npass = ax.rad.npass;
                    for (int i = 0; i < kMaxPass; ++i) op.radix[i] = i < ax.rad.npass ? ax.rad.r[i] : 1;
                    ops.push_back(op);
                    maxl = std::max(maxl, ax.nlines);
                }
                ln.count = int(ops.size()) - ln.first;
                ln.P = lines_P(ax.L);
                ln.lds = size_t(2) * ln.P * ax.L * sizeof(cpx<T>);
                ln.gx = (maxl + 2 * ln.P - 1) / (2 * ln.P);
                if (ln.count > 0) launches_.push_back(ln);
            }
        }
        RETCHK(alloc(&ops_dev_, std::max<size_t>(1, ops.size())));
        HIPCHK(hipMemcpy(ops_dev_, ops.data(), ops.size() * sizeof(LineOp<T>), hipMemcpyHostToDevice));

        // ---- prep: fields whose Boltzmann factor the chains use
        std::memset(&prep_, 0, sizeof(prep_));
        bool used[13] = {false};
        for (auto &ch : chains_) used[ch.fa] = used[ch.fb] = true;
        for (int f = 0; f < 13; ++f)
            if (used[f]) prep_.fields[prep_.nfields++] = f;
        prep_.w = c.w;
        prep_.ds = DS;
        prep_.V = c.V;
        prep_.nx = c.NX; prep_.ny = c.NY; prep_.nz = c.NZ;
        prep_.delta = c.p.seed_area_delta;

        // ---- Q slots: dbk1/2 from q-dagger, dbk3/4 and the rest from q
        // (SCFT.c:1271, 1314, 1357, 1399, 1422)
        std::vector<const T *> ends;
        FinArgs fin{};
        for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
            const int ns = d.ns_sp[sp];
            if (ns <= 0) continue;
            const T *src = (sp < 2) ? qd_[sp] : q_[sp];
            ends.push_back(src + (size_t)(ns - 1) * V);
            fin.qsp[fin.nq] = sp;
            double coef = 0.0;
            if (sp < 4) coef = c.p.vol_dbk[sp] / d.alpha[sp] * (1.0 - d.coating_vol);
            else if (sp < 14) coef = c.p.vol_homo[sp - 4] / d.alpha[sp] * (1.0 - d.coating_vol);
            else coef = c.p.mask ? c.p.vol_s[sp - 14] / d.alpha[sp] : 0.0;
            fin.lnq_coef[fin.nq] = coef;
            fin.nq++;
        }
        nq_ = fin.nq;
        RETCHK(alloc(&ends_dev_, std::max<size_t>(1, ends.size())));
        HIPCHK(hipMemcpy(ends_dev_, ends.data(), ends.size() * sizeof(T *), hipMemcpyHostToDevice));
        fin.V = c.V;
        fin.nblocks = c.nred;
        fin.nact = d.n_active;
        fin.qpart = c.qpart;
        fin.dpart = c.dpart;
        fin.upart = c.upart;
        fin.out = c.scal;
        HIPCHK(hipMemcpy(c.fin, &fin, sizeof(fin), hipMemcpyHostToDevice));

        // ---- density terms (SCFT.c:1092-1513)
        DenArgs<T> den;
        std::memset(&den, 0, sizeof(den));
        auto term = [&](int sp, int start, int stop, int field, double frac) {
            DenTerm<T> &t = den.term[den.nterms++];
            t.q = q_[sp];
            t.qd = qd_[sp];
            t.ns = d.ns_sp[sp];
            t.start = start;
            t.stop = stop;
            t.field = field;
            t.qslot = sp;
            t.frac = frac;
        };
        const double cv = d.coating_vol;
        for (int h = 0; h < 10; ++h)
            if (d.ns_sp[4 + h] > 0) term(4 + h, 0, d.ns_sp[4 + h] - 1, h, c.p.vol_homo[h] * (1.0 - cv));
        for (int b = 0; b < 4; ++b) {
            if (d.ns_sp[b] <= 0) continue;
            const double fr = c.p.vol_dbk[b] * (1.0 - cv);
            term(b, 0, d.ns_x[b] - 1, kDbkField[b][0], fr);
            term(b, d.ns_x[b] - 1, d.ns_sp[b] - 1, kDbkField[b][1], fr);
        }
        for (int s = 0; s < 3; ++s)
            if (d.ns_sp[14 + s] > 0) term(14 + s, 0, d.ns_sp[14 + s] - 1, 10 + s, c.p.mask ? c.p.vol_s[s] : 0.0);
        bool carries[13] = {false};
        for (int i = 0; i < den.nterms; ++i) carries[den.term[i].field] = true;
        for (int f = 0; f < 13; ++f)
            if (carries[f]) den.phif[den.nphi++] = f;
        for (int f = 0; f < 13; ++f)
            for (int g = 0; g < 13; ++g) den.chi[f][g] = d.chi_eff[f][g];
        den.V = c.V;
        den.mask = c.p.mask;
        den.Q = c.scal->Q;
        den.w = c.w;
        den.p = c.pf;
        den.tot = c.tot;
        den.phitot = c.phitot;
        den.partial = c.dpart;
        RETCHK(alloc(&den_dev_, 1));
        HIPCHK(hipMemcpy(den_dev_, &den, sizeof(den), hipMemcpyHostToDevice));
        return SCFT_OK;
    }

    int propagate(unsigned long iter) override
    {
        scft_ctx &c = *c_;
        PrepArgs pa = prep_;
        pa.seed_mode = 0;
        if (c.d.ns_sp[0] > 0 && c.p.enable_seed)
            pa.seed_mode = (iter <= (unsigned long)c.p.seed_steps) ? 1 : 2;
        hipLaunchKernelGGL(k_prep<T>, dim3(grid_for(c.V)), dim3(kThreads), 0, c.stream, pa, E_,
                           c.d.ns_sp[0] > 0 ? q_[0] : nullptr);
        for (const Launch &ln : launches_) {
            hipLaunchKernelGGL(k_lineop<T>, dim3(ln.gx, ln.count), dim3(kThreads), ln.lds, c.stream,
                               ops_dev_ + ln.first, ln.P);
        }
        HIPCHK(hipGetLastError());
        return SCFT_OK;
    }

    int densities() override
    {
        scft_ctx &c = *c_;
        if (nq_ > 0) {
            hipLaunchKernelGGL(k_qsum<T>, dim3(c.nred, nq_), dim3(kThreads), 0, c.stream,
                               (const T *const *)ends_dev_, c.V, c.qpart);
            hipLaunchKernelGGL(k_qfinal, dim3(nq_), dim3(kThreads), 0, c.stream, (const FinArgs *)c.fin);
        }
        hipLaunchKernelGGL(k_density<T>, dim3(c.nred), dim3(kThreads), 0, c.stream,
                           (const DenArgs<T> *)den_dev_);
        HIPCHK(hipGetLastError());
        return SCFT_OK;
    }

    int update(float mf, const double *noise_dev) override
    {
        scft_ctx &c = *c_;
        UpdArgs u = c.upd_host;
        u.amp = c.p.dt_noise * mf;
        u.noise = (mf != 0.0f) ? noise_dev : nullptr;
        HIPCHK(hipMemcpyAsync(c.upd, &u, sizeof(u), hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_update, dim3(c.nred), dim3(kThreads), 0, c.stream, (const UpdArgs *)c.upd);
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, c.stream, (const FinArgs *)c.fin);
        hipLaunchKernelGGL(k_pshift, dim3(grid_for(c.V)), dim3(kThreads), 0, c.stream, c.pf, c.V,
                           (const Scalars *)c.scal);
        HIPCHK(hipGetLastError());
        return SCFT_OK;
    }

    int set_mask(const double *mask_dev) override
    {
        scft_ctx &c = *c_;
        for (int s = 0; s < 3; ++s)
            if (c.d.ns_sp[14 + s] > 0)
                hipLaunchKernelGGL(k_convert<T>, dim3(grid_for(c.V)), dim3(kThreads), 0, c.stream,
                                   mask_dev + (size_t)s * c.V, q_[14 + s], c.V);
        HIPCHK(hipGetLastError());
        return SCFT_OK;
    }

    int get_propagator(int sp, int which, int s, double *host) override
    {
        scft_ctx &c = *c_;
        if (sp < 0 || sp >= SCFT_NSPECIES || c.d.ns_sp[sp] <= 0 || s < 0 || s >= c.d.ns_sp[sp]) {
            set_err("scft_get_propagator: no such slice");
            return SCFT_EINVAL;
        }
        const T *src = (which ? qd_[sp] : q_[sp]) + (size_t)s * c.V;
        double *tmp = nullptr;
        RETCHK(dev_alloc(&tmp, c.V));
        hipLaunchKernelGGL(k_widen<T>, dim3(grid_for(c.V)), dim3(kThreads), 0, c.stream, src, tmp, c.V);
        hipError_t e = hipMemcpyAsync(host, tmp, sizeof(double) * c.V, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) {
            set_err(std::string("get_propagator: ") + hipGetErrorString(e));
            return SCFT_EHIP;
        }
        return SCFT_OK;
    }

private:
    struct Chain {
        T *base;
        T *tmp;
        int nsteps, fa, fb, switch_at;
    };
    struct Launch {
        int first, count, P, gx;
        size_t lds;
    };

    template <typename X>
    int alloc(X **p, size_t n)
    {
        RETCHK(dev_alloc(p, n));
        allocs_.push_back(*p);
        return SCFT_OK;
    }

    void fill(T *p, T v, size_t n)
    {
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(kThreads), 0, c_->stream, p, v, n);
    }

    void add_chain(T *base, int nsteps, int fa, int fb, int switch_at)
    {
        chains_.push_back(Chain{base, nullptr, nsteps, fa, fb, switch_at});
    }

    static int lines_P(int L)
    {
        const int budget = 64 * 1024;
        int P = budget / int(2 * L * sizeof(cpx<T>));
        return std::max(1, std::min(P, 16));
    }

    int setup_axis(Axis &a, double alpha)
    {
        if (!plan_radices(a.L, a.rad)) {
            set_err("transform length " + std::to_string(a.L) + " has too many factors");
            return SCFT_EUNSUPPORTED;
        }
        if (size_t(2) * a.L * sizeof(cpx<T>) > 64 * 1024) {
            set_err("transform length " + std::to_string(a.L) + " exceeds the LDS tile");
            return SCFT_EUNSUPPORTED;
        }
        std::vector<double> k = axis_k(a.L, alpha);
        std::vector<T> kt(k.begin(), k.end());
        std::vector<cpx<T>> tw = twiddles<T>(a.L);
        T *Kd = nullptr;
        cpx<T> *twd = nullptr;
        RETCHK(alloc(&Kd, a.L));
        RETCHK(alloc(&twd, a.L));
        HIPCHK(hipMemcpy(Kd, kt.data(), a.L * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(twd, tw.data(), a.L * sizeof(cpx<T>), hipMemcpyHostToDevice));
        a.K = Kd;
        a.tw = twd;
        return SCFT_OK;
    }

    scft_ctx *c_;
    std::vector<void *> allocs_;
    T *E_ = nullptr;
    T *q_[SCFT_NSPECIES] = {};
    T *qd_[SCFT_NSPECIES] = {};
    std::vector<Chain> chains_;
    std::vector<Axis> axes_;
    std::vector<Launch> launches_;
    int maxsteps_ = 0;
    LineOp<T> *ops_dev_ = nullptr;
    T **ends_dev_ = nullptr;
    int nq_ = 0;
    DenArgs<T> *den_dev_ = nullptr;
    PrepArgs prep_{};
};

void free_ctx(scft_ctx *c)
{
    if (!c) return;
    c->eng.reset();
    double *bufs[] = {c->w, c->pf, c->tot, c->phitot, c->mask, c->qpart, c->dpart, c->upart, c->noise};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (c->scal) (void)hipFree(c->scal);
    if (c->fin) (void)hipFree(c->fin);
    if (c->upd) (void)hipFree(c->upd);
    if (c->scal_host) (void)hipHostFree(c->scal_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int fetch_out(scft_ctx *c, float mf, scft_iter_out *out)
{
    HIPCHK(hipMemcpyAsync(c->scal_host, c->scal, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!out) return SCFT_OK;
    std::memset(out, 0, sizeof(*out));
    const Scalars &s = *c->scal_host;
    out->free_energy = s.F;
    out->rmsdev = s.rms;
    out->dw_rms = s.dw_rms;
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
        if (c->d.ns_sp[sp] <= 0) continue;
        out->qsum[sp] = s.Q[sp];
        const double qmod = std::pow(s.Q[sp], 2);
        out->chempot[sp] = (qmod == 0) ? 0.0 : -std::log(qmod) / 2.0;
    }
    out->mf = mf;
    return SCFT_OK;
}

}  // namespace

extern "C" {

const char *scft_last_error(void) { return g_err.c_str(); }

int scft_create(const scft_params *p, int precision, scft_ctx **out)
{
    if (!p || !out || (precision != SCFT_FP64 && precision != SCFT_FP32_SPLIT)) {
        set_err("scft_create: bad argument");
        return SCFT_EINVAL;
    }
    *out = nullptr;
    scft_ctx *c = new scft_ctx();
    c->p = *p;
    c->precision = precision;
    int rc = scft_derive(p, &c->d);
    if (rc != SCFT_OK) {
        set_err(rc == SCFT_EVOLUME ? "volumes do not sum to 1" : "invalid parameters");
        delete c;
        return rc;
    }
    c->NX = c->d.nx;
    c->NY = c->d.ny;
    c->NZ = c->d.nz;
    const size_t V = size_t(c->NX) * c->NY * c->NZ;
    if (V > size_t(1) << 30) {
        set_err("grid too large");
        delete c;
        return SCFT_EUNSUPPORTED;
    }
    c->V = int(V);
    c->nred = grid_for(V, kRedBlocks);
    auto fail = [&](int r) {
        free_ctx(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        return fail(SCFT_EHIP);
    }
    if ((rc = dev_alloc(&c->w, 13 * V)) || (rc = dev_alloc(&c->pf, V)) || (rc = dev_alloc(&c->tot, 13 * V)) ||
        (rc = dev_alloc(&c->phitot, V)) || (rc = dev_alloc(&c->mask, 3 * V)) ||
        (rc = dev_alloc(&c->qpart, size_t(SCFT_NSPECIES) * c->nred)) || (rc = dev_alloc(&c->dpart, 2 * size_t(c->nred))) ||
        (rc = dev_alloc(&c->upart, 2 * size_t(c->nred))) || (rc = dev_alloc(&c->scal, 1)) ||
        (rc = dev_alloc(&c->fin, 1)) || (rc = dev_alloc(&c->upd, 1)))
        return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void **>(&c->scal_host), sizeof(Scalars), hipHostMallocDefault) != hipSuccess) {
        set_err("hipHostMalloc failed");
        return fail(SCFT_ENOMEM);
    }
    // zero state: fields, densities (absent chemistries stay 0, SCFT.c:785-843)
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(13 * V)), dim3(kThreads), 0, c->stream, c->w, 0.0, 13 * V);
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(V)), dim3(kThreads), 0, c->stream, c->pf, 0.0, V);
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(13 * V)), dim3(kThreads), 0, c->stream, c->tot, 0.0, 13 * V);
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(V)), dim3(kThreads), 0, c->stream, c->phitot, 0.0, V);
    hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(3 * V)), dim3(kThreads), 0, c->stream, c->mask, 0.0, 3 * V);
    hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(kThreads), 0, c->stream, reinterpret_cast<double *>(c->scal), 0.0,
                       sizeof(Scalars) / sizeof(double));

    // update arguments (constant per context)
    UpdArgs &u = c->upd_host;
    std::memset(&u, 0, sizeof(u));
    u.V = c->V;
    u.nx = c->NX;
    u.ny = c->NY;
    u.nz = c->NZ;
    for (int f = 0; f < 13; ++f)
        if (c->d.active[f]) u.act[u.nact++] = f;
    for (int f = 0; f < 13; ++f) {
        u.phibar[f] = c->d.phi_avg[f];
        for (int g = 0; g < 13; ++g) u.chi[f][g] = c->d.chi_eff[f][g];
    }
    u.g1 = p->g1;
    u.g2 = p->g1 / 6.0;
    u.w = c->w;
    u.p = c->pf;
    u.tot = c->tot;
    u.partial = c->upart;

    if (precision == SCFT_FP64) {
        auto e = std::make_unique<Engine<double>>(c);
        rc = e->init();
        c->eng = std::move(e);
    } else {
        auto e = std::make_unique<Engine<float>>(c);
        rc = e->init();
        c->eng = std::move(e);
    }
    if (rc != SCFT_OK) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_err("initialisation failed");
        return fail(SCFT_EHIP);
    }
    *out = c;
    return SCFT_OK;
}

int scft_destroy(scft_ctx *ctx)
{
    if (!ctx) return SCFT_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
    return SCFT_OK;
}

int scft_get_derived(const scft_ctx *ctx, scft_derived *out)
{
    if (!ctx || !out) return SCFT_EINVAL;
    *out = ctx->d;
    return SCFT_OK;
}

int scft_set_fields(scft_ctx *c, const double *w, const double *p)
{
    if (!c || !w || !p) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(c->w, w, sizeof(double) * 13 * c->V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->pf, p, sizeof(double) * c->V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_get_fields(scft_ctx *c, double *w, double *p)
{
    if (!c || !w || !p) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(w, c->w, sizeof(double) * 13 * c->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(p, c->pf, sizeof(double) * c->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_set_mask(scft_ctx *c, const double *mask)
{
    if (!c || !mask) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(c->mask, mask, sizeof(double) * 3 * c->V, hipMemcpyHostToDevice, c->stream));
    RETCHK(c->eng->set_mask(c->mask));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_get_densities(scft_ctx *c, double *phi, double *phi_tot)
{
    if (!c || !phi) return SCFT_EINVAL;
    HIPCHK(hipMemcpyAsync(phi, c->tot, sizeof(double) * 13 * c->V, hipMemcpyDeviceToHost, c->stream));
    if (phi_tot)
        HIPCHK(hipMemcpyAsync(phi_tot, c->phitot, sizeof(double) * c->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

int scft_get_propagator(scft_ctx *c, int sp, int which, int s, double *out)
{
    if (!c || !out) return SCFT_EINVAL;
    return c->eng->get_propagator(sp, which, s, out);
}

int scft_propagate(scft_ctx *c, unsigned long iter)
{
    if (!c) return SCFT_EINVAL;
    return c->eng->propagate(iter);
}

int scft_densities(scft_ctx *c, double *qsum, double *chempot)
{
    if (!c) return SCFT_EINVAL;
    RETCHK(c->eng->densities());
    if (qsum || chempot) {
        scft_iter_out o;
        RETCHK(fetch_out(c, 0.0f, &o));
        for (int sp = 0; sp < SCFT_NSPECIES; ++sp) {
            if (qsum) qsum[sp] = o.qsum[sp];
            if (chempot) chempot[sp] = o.chempot[sp];
        }
    }
    return SCFT_OK;
}

static int upload_noise(scft_ctx *c, float mf, const double *noise, const double **dev)
{
    *dev = nullptr;
    if (mf == 0.0f) return SCFT_OK;
    if (!noise) {
        set_err("noise is required when mf != 0");
        return SCFT_EINVAL;
    }
    const size_t n = size_t(c->d.n_active) * c->V;
    if (!c->noise) RETCHK(dev_alloc(&c->noise, n));
    HIPCHK(hipMemcpyAsync(c->noise, noise, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    *dev = c->noise;
    return SCFT_OK;
}

int scft_update_fields(scft_ctx *c, unsigned long iter, float mf, const double *noise, scft_iter_out *out)
{
    (void)iter;
    if (!c) return SCFT_EINVAL;
    const double *nd = nullptr;
    RETCHK(upload_noise(c, mf, noise, &nd));
    RETCHK(c->eng->update(mf, nd));
    return fetch_out(c, mf, out);
}

int scft_iterate(scft_ctx *c, unsigned long iter, float mf, const double *noise, scft_iter_out *out)
{
    if (!c) return SCFT_EINVAL;
    const double *nd = nullptr;
    RETCHK(upload_noise(c, mf, noise, &nd));
    RETCHK(c->eng->propagate(iter));
    RETCHK(c->eng->densities());
    RETCHK(c->eng->update(mf, nd));
    return fetch_out(c, mf, out);
}

int scft_sync(scft_ctx *c)
{
    if (!c) return SCFT_EINVAL;
    HIPCHK(hipStreamSynchronize(c->stream));
    return SCFT_OK;
}

}  // extern "C"
