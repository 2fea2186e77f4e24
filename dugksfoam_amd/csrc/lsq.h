// lsq.h -- least-squares gradient vectors of a polyMesh (host code shared by the step engine and
// the adjoint).
//
// OpenFOAM's leastSquaresVectors as the reference uses it (zeroBoundaryVectors.C:90-230 with the
// non-coupled boundary contributions of lines 154-166 enabled):
//   dd_c  = sum over internal faces (|Sf|/|d|^2) d d, weighted (1 - w) on the owner and w on the
//           neighbour, plus (|Sb|/|db|^2) db db over boundary faces ((1 - w) times that on cyclic
//           faces);
//   pVec_f = (1 - w) (|Sf|/|d|^2) dd_own^-1 d,  nVec_f = -w (|Sf|/|d|^2) dd_nei^-1 d,
//   pVec_b = (|Sb|/|db|^2) dd^-1 db  (cyclic: (1 - w) times that),
// where components of dd that vanish on the first cell (the empty direction of a 2-D mesh) are set
// to 1 before the inversion and removed after it (inv(symmTensorField)).
#pragma once

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/dugks.h"

namespace dugks_lsq {

inline double mag3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// inverse of a symmetric tensor (xx, xy, xz, yy, yz, zz) by cofactors
inline void symm_inverse(const double *s, double *r)
{
    double det = s[0] * s[3] * s[5] + s[1] * s[4] * s[2] + s[2] * s[1] * s[4] - s[0] * s[4] * s[4] -
                 s[1] * s[1] * s[5] - s[2] * s[3] * s[2];
    r[0] = (s[3] * s[5] - s[4] * s[4]) / det;
    r[1] = (s[2] * s[4] - s[1] * s[5]) / det;
    r[2] = (s[1] * s[4] - s[2] * s[3]) / det;
    r[3] = (s[0] * s[5] - s[2] * s[2]) / det;
    r[4] = (s[1] * s[2] - s[0] * s[4]) / det;
    r[5] = (s[0] * s[3] - s[1] * s[1]) / det;
}

// pvec[3 n_faces], nvec[3 n_internal_faces].  bkind[b] is the patch type of boundary face b.
// Returns nullptr on success, else the reason the vectors do not exist.
inline const char *vectors(const dugks_case *cs, const std::vector<int> &bkind, std::vector<double> &pvec,
                           std::vector<double> &nvec)
{
    const int nc = cs->n_cells, nfi = cs->n_internal_faces, nf = cs->n_faces, nb = nf - nfi;
    std::vector<double> dd(6 * (size_t)nc, 0.0), magSf(nf);
    pvec.assign(3 * (size_t)nf, 0.0);
    nvec.assign(3 * (size_t)nfi, 0.0);
    for (int f = 0; f < nf; ++f) magSf[f] = mag3(cs->Sf + 3 * f);
    // wdd = (|Sf|/|d|^2) d d; dd_own += (1 - w) wdd; dd_nei += w wdd (leastSquaresVectors order)
    auto outer_add = [&](int cell, double m, const double *d) {
        double *t = &dd[6 * (size_t)cell];
        t[0] += m * (d[0] * d[0]); t[1] += m * (d[0] * d[1]); t[2] += m * (d[0] * d[2]);
        t[3] += m * (d[1] * d[1]); t[4] += m * (d[1] * d[2]); t[5] += m * (d[2] * d[2]);
    };
    const double *Cc = cs->C, *w = cs->weights;
    for (int f = 0; f < nfi; ++f) {
        int o = cs->owner[f], n = cs->neighbour[f];
        double d[3] = {Cc[3 * n] - Cc[3 * o], Cc[3 * n + 1] - Cc[3 * o + 1], Cc[3 * n + 2] - Cc[3 * o + 2]};
        double m = magSf[f] / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double wdd[6] = {m * (d[0] * d[0]), m * (d[0] * d[1]), m * (d[0] * d[2]),
                               m * (d[1] * d[1]), m * (d[1] * d[2]), m * (d[2] * d[2])};
        double *to = &dd[6 * (size_t)o], *tn = &dd[6 * (size_t)n];
        for (int i = 0; i < 6; ++i) {
            to[i] += (1 - w[f]) * wdd[i];
            tn[i] += w[f] * wdd[i];
        }
    }
    for (int b = 0; b < nb; ++b) {
        int f = nfi + b;
        const double *d = cs->delta + 3 * b;
        double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(d2 > 0)) return "zero boundary delta";
        double m = bkind[b] == DUGKS_PATCH_CYCLIC ? (1 - w[f]) * magSf[f] / d2 : magSf[f] / d2;
        outer_add(cs->owner[f], m, d);
    }
    std::vector<double> inv(6 * (size_t)nc);
    {
        const double *s0 = &dd[0];
        double scale = s0[0] * s0[0] + 2 * s0[1] * s0[1] + 2 * s0[2] * s0[2] + s0[3] * s0[3] + 2 * s0[4] * s0[4] +
                       s0[5] * s0[5];
        bool rm[3] = {s0[0] * s0[0] / scale < 1e-15, s0[3] * s0[3] / scale < 1e-15, s0[5] * s0[5] / scale < 1e-15};
        for (int c = 0; c < nc; ++c) {
            double s[6];
            std::memcpy(s, &dd[6 * (size_t)c], sizeof s);
            if (rm[0]) s[0] += 1;
            if (rm[1]) s[3] += 1;
            if (rm[2]) s[5] += 1;
            double *r = &inv[6 * (size_t)c];
            symm_inverse(s, r);
            if (rm[0]) r[0] -= 1;
            if (rm[1]) r[3] -= 1;
            if (rm[2]) r[5] -= 1;
            for (int i = 0; i < 6; ++i)
                if (!std::isfinite(r[i])) return "singular least-squares tensor";
        }
    }
    auto sdot = [](const double *s, const double *d, double *r) {
        r[0] = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
        r[1] = s[1] * d[0] + s[3] * d[1] + s[4] * d[2];
        r[2] = s[2] * d[0] + s[4] * d[1] + s[5] * d[2];
    };
    for (int f = 0; f < nfi; ++f) {
        int o = cs->owner[f], n = cs->neighbour[f];
        double d[3] = {Cc[3 * n] - Cc[3 * o], Cc[3 * n + 1] - Cc[3 * o + 1], Cc[3 * n + 2] - Cc[3 * o + 2]};
        double m = magSf[f] / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double a[3], bb[3];
        sdot(&inv[6 * (size_t)o], d, a);
        sdot(&inv[6 * (size_t)n], d, bb);
        for (int i = 0; i < 3; ++i) {
            pvec[3 * f + i] = (1 - w[f]) * m * a[i];
            nvec[3 * f + i] = -w[f] * m * bb[i];
        }
    }
    for (int b = 0; b < nb; ++b) {
        int f = nfi + b;
        const double *d = cs->delta + 3 * b;
        double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double a[3];
        sdot(&inv[6 * (size_t)cs->owner[f]], d, a);
        double m = bkind[b] == DUGKS_PATCH_CYCLIC ? ((1 - w[f]) * magSf[f] / d2) : magSf[f] * (1.0 / d2);
        for (int i = 0; i < 3; ++i) pvec[3 * f + i] = m * a[i];
    }
    return nullptr;
}

}  // namespace dugks_lsq
