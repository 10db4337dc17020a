// scft_setup.cpp — host-side setup derivation of libscft.
//
// Replaces the prologue of SCFT.c's main(): the volume check (SCFT.c:84-101),
// grid sizing (103-105), the chi_eff copy and zeroing (107-119, 303-396),
// contour discretisation per species (121-288), mean volume fractions
// (290-301) with the coating rescale (718-734), and the active-field rule of
// funUpdate (1993-2042).
#include <cmath>
#include <cstring>

#include "../../include/scft.h"

namespace {

int grid_points(double L, double dr)
{
    // SCFT.c:103-105: a length of exactly 1.0 collapses the axis.
    if (L == 1.0) return 1;
    return static_cast<int>((L / dr) + 0.5);
}

int odd_down(int n) { return (n % 2 == 0) ? n - 1 : n; }

}  // namespace

extern "C" int scft_derive(const scft_params *p, scft_derived *d)
{
    if (!p || !d) return SCFT_EINVAL;
    std::memset(d, 0, sizeof(*d));
    const double eps = 1e-6;

    double vsum = 0.0;
    for (int b = 0; b < 4; ++b) vsum += p->vol_dbk[b];
    for (int h = 0; h < 10; ++h) vsum += p->vol_homo[h];
    if (std::fabs(1.0 - vsum) >= eps) return SCFT_EVOLUME;
    if (p->ns < 3 || p->dr_target <= 0.0) return SCFT_EINVAL;

    d->nx = grid_points(p->lx, p->dr_target);
    d->ny = grid_points(p->ly, p->dr_target);
    d->nz = grid_points(p->lz, p->dr_target);
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return SCFT_EINVAL;
    d->nzext = p->mask ? 2 * (d->nz + 1) : d->nz;

    const int NS = p->ns;
    const double DS = 1.0 / (NS - 1);
    for (int b = 0; b < 4; ++b) {
        if (!(p->vol_dbk[b] > 0.0)) continue;
        const int n = odd_down(static_cast<int>(std::floor(p->len_dbk[b] * NS)));
        const int nx = odd_down(static_cast<int>(std::floor(p->f_dbk[b] * n)) + 1);
        d->ns_sp[b] = n;
        d->ns_x[b] = nx;
        d->ns_y[b] = n - nx;
        d->alpha[b] = static_cast<double>((n - 1) * DS);
    }
    for (int h = 0; h < 10; ++h) {
        if (!(p->vol_homo[h] > 0.0)) continue;
        int n = static_cast<int>(std::floor(p->len_homo[h] * NS));
        if (n % 2 == 0 || n > NS) n -= 1;   // homopolymers are also capped at NS
        d->ns_sp[4 + h] = n;
        d->alpha[4 + h] = static_cast<double>((n - 1) * DS);
    }
    if (p->mask) {
        for (int s = 0; s < 3; ++s) {
            if (!(p->vol_s[s] > 0.0)) continue;
            const int n = odd_down(static_cast<int>(std::floor(p->len_s[s] * NS)));
            d->ns_sp[14 + s] = n;
            d->alpha[14 + s] = static_cast<double>((n - 1) * DS);
        }
    }
    // Every present species needs at least one Simpson interval.
    for (int sp = 0; sp < SCFT_NSPECIES; ++sp)
        if (d->ns_sp[sp] != 0 && d->ns_sp[sp] < 3) return SCFT_EINVAL;
    for (int b = 0; b < 4; ++b)
        if (d->ns_sp[b] && (d->ns_x[b] < 1 || d->ns_y[b] < 1)) return SCFT_EINVAL;

    // Mean volume fractions: a diblock splits its volume f : 1-f between
    // its two chemistries (pairs A/B, D/E, F/G, H/I).
    static const int kPair[4][2] = {{0, 1}, {3, 4}, {5, 6}, {7, 8}};
    double *pa = d->phi_avg;
    for (int f = 0; f < 10; ++f) pa[f] = p->vol_homo[f];
    for (int b = 0; b < 4; ++b) {
        pa[kPair[b][0]] += p->f_dbk[b] * p->vol_dbk[b];
        pa[kPair[b][1]] += (1.0 - p->f_dbk[b]) * p->vol_dbk[b];
    }

    bool zero_field[SCFT_NFIELDS] = {false};
    for (int f = 0; f < 10; ++f) zero_field[f] = std::fabs(pa[f]) <= eps;
    for (int s = 0; s < 3; ++s) {
        if (p->mask && !(std::fabs(p->vol_s[s]) <= eps)) {
            pa[10 + s] = p->vol_s[s];
        } else {
            zero_field[10 + s] = true;
            pa[10 + s] = 0.0;
        }
    }
    for (int f = 0; f < SCFT_NFIELDS; ++f)
        for (int g = 0; g < SCFT_NFIELDS; ++g)
            d->chi_eff[f][g] = (f == g || zero_field[f] || zero_field[g]) ? 0.0 : p->chi[f][g];

    if (p->mask) {
        d->coating_vol = p->vol_s[0] + p->vol_s[1] + p->vol_s[2];
        for (int f = 0; f < 10; ++f) pa[f] = pa[f] * (1.0 - d->coating_vol);
    }

    // Active fields: own homopolymer/brush volume > 0 (brushes count even
    // with MASK 0), or the diblock that carries the field is present.
    for (int f = 0; f < 10; ++f) d->active[f] = p->vol_homo[f] > 0.0;
    for (int s = 0; s < 3; ++s) d->active[10 + s] = p->vol_s[s] > 0.0;
    for (int b = 0; b < 4; ++b)
        if (p->vol_dbk[b] > 0.0) d->active[kPair[b][0]] = d->active[kPair[b][1]] = 1;
    d->n_active = 0;
    for (int f = 0; f < SCFT_NFIELDS; ++f) d->n_active += d->active[f] ? 1 : 0;
    return SCFT_OK;
}
