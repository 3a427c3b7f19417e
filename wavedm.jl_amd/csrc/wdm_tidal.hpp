// Milky Way tidal potential evaluated in the kick (opt-in; V_TIDAL).
//
// src/physics/tidal_fields.jl:64-146, MW_tidal_interpolate branch, as the code
// intends it (in the reference the result never reaches the kick: its last
// line `Φ_all += pot_tidal` rebinds a local, :117):
//   pot(x)  = interpolate(MW_grid, MW_Phi, x + o)      o: trajectory offset of the kick
//   pot    -= pot[c]                                    c: density peak (:114)
//   pot    -= sum_a g_a (i_a - div(N_a, 2)),            g_a = (pot[c+e_a] - pot[c-e_a]) / 2,
//             i_a 1-based                               (cancel_field_gradient_at_center!, :126-140)
//   pot    += sum(ramps * rho) / sum(rho)               (:141-143)
// The host sets o, the centre value c0, the g_a and the constant per kick
// (TidalModel); the kernel interpolates the uploaded table at every cell.
// Interpolation: multilinear on the uniform grid lo + i (hi - lo)/(n - 1),
// coordinates clamped into it (the GridInterpolations rule is not in this
// image; the oracle states the same rule).
#pragma once

#include "wdm_math.hpp"

namespace wdm {

struct TidalModel {
    const double* table;  // n[0]*n[1]*n[2] FP64, x fastest
    int n[3];
    double lo[3], step[3], inv[3];
    double off[3];        // trajectory offset of this kick
    double c0;            // interpolated pot at the centre cell (subtracted)
    double g[3];          // central differences at the centre, index units
    double pe;            // sum(ramps * rho) / sum(rho)
    int half[3];          // div(N_a, 2) of the simulation grid
};

// multilinear interpolation of the table at p (clamped into the grid)
WDM_HDFI double tidal_interp(const TidalModel& t, double px, double py, double pz) {
    const double p[3] = {px, py, pz};
    int i0[3];
    double w[3];
    for (int a = 0; a < 3; ++a) {
        const double hi = t.lo[a] + t.step[a] * double(t.n[a] - 1);
        const double c = p[a] < t.lo[a] ? t.lo[a] : (p[a] > hi ? hi : p[a]);
        const double f = (c - t.lo[a]) * t.inv[a];
        int i = int(f);
        if (i > t.n[a] - 2) i = t.n[a] - 2;
        if (i < 0) i = 0;
        i0[a] = i;
        w[a] = f - double(i);
    }
    const size_t sx = 1, sy = size_t(t.n[0]), sz = size_t(t.n[0]) * size_t(t.n[1]);
    const double* b = t.table + size_t(i0[0]) * sx + size_t(i0[1]) * sy + size_t(i0[2]) * sz;
    const double c00 = b[0] * (1.0 - w[0]) + b[sx] * w[0];
    const double c10 = b[sy] * (1.0 - w[0]) + b[sy + sx] * w[0];
    const double c01 = b[sz] * (1.0 - w[0]) + b[sz + sx] * w[0];
    const double c11 = b[sz + sy] * (1.0 - w[0]) + b[sz + sy + sx] * w[0];
    const double c0 = c00 * (1.0 - w[1]) + c10 * w[1];
    const double c1 = c01 * (1.0 - w[1]) + c11 * w[1];
    return c0 * (1.0 - w[2]) + c1 * w[2];
}

// the tidal potential the kick adds at cell (i, j, k) with coordinates (x, y, z)
WDM_HDFI double tidal_V(const TidalModel& t, double x, double y, double z, int i, int j, int k) {
    const double p = tidal_interp(t, x + t.off[0], y + t.off[1], z + t.off[2]);
    const double ramp = t.g[0] * double(i + 1 - t.half[0]) + t.g[1] * double(j + 1 - t.half[1]) +
                        t.g[2] * double(k + 1 - t.half[2]);
    return p - t.c0 - ramp + t.pe;
}

}  // namespace wdm
