// dugks_engine.hip -- MI355X (gfx950) DUGKS time-step engine behind the C-ABI of include/dugks.h.
//
// One time step of dugksFoam's fvDVM::evolution() (src/fvDVM/fvDVM/fvDVM.C:1086-1108) as two
// kernels over resident device state:
//
//   k_recon  (one workgroup per cell, all velocities):
//            S2 least-squares gradient of the resident gBarPvol and upwind face reconstruction;
//            each face value is written once, by the upwind cell.  Also the lagged fixedGradient
//            boundary gradient of the next step.
//   k_fmom   (one workgroup per face, all velocities):
//            S3/S4 Maxwell wall (outGoing, rho_w, incoming Maxwellian), S6 face moments (rho, rhoU,
//            rhoE and the heat flux) reduced across the velocity lanes; face macros, equilibrium
//            constants and exp tables.
//   k_cells  (one workgroup per cell, all velocities):
//            S7 face relaxation, S8 flux divergence and g~ update in the reference's per-cell face
//            order, S9 cell moments and macros, then S1 of the next step (gBarPvol = (1-r) g~ + r geq),
//            written to the other ping-pong buffer.
//
// Data layout: distributions are [cell][velocity] with the velocity index contiguous and padded to
// NVP = NT*NITER; a workgroup's lanes are velocities of one cell, so every mesh quantity is
// wave-uniform (scalar loads) and every distribution access is a coalesced row segment.  The
// resident state is gBarPvol (g-bar-plus) and hBarPvol; g~ is recovered as (gbar - r geq)/(1 - r)
// with the same r and geq that formed gbar.  Velocities are stored grouped by the sign pattern of
// (xi_x, xi_y, xi_z) so that the upwind side of an axis-aligned face is uniform across a wavefront.
//
// Face heat flux: q_f = sum 1/2 w c (|c|^2 g + h), c = xi - U_f, needs U_f, itself a velocity sum.
// Both come from one pass by accumulating moments about a reference velocity Uh (the face
// velocity of the previous step; the wall velocity on walls) and shifting exactly:
//   q = Q - 2 P d + |d|^2 B - d A + 2 d (B.d) - d |d|^2 M,  d = U - Uh,
// with Q_i = sum 1/2 w ch_i (|ch|^2 g + h), P_ij = sum 1/2 w ch_i ch_j g, ch = xi - Uh, and
// B, A, M from the conserved sums.  The cell heat flux uses the same identity about the cell
// velocity of the previous step.
//
// Equilibria: exp(-|U - xi|^2 / 2RT) is evaluated as the product of per-axis factors of the
// tensor-product velocity grid, tabulated once per cell / face / wall face.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dugks.h"

namespace {

constexpr double VSMALL = 1.0e-300;   // OpenFOAM VSMALL (doubleScalar.H)
constexpr double PI = 3.14159265358979323846;
constexpr int KIND_OWN = 0;   // internal face, this cell is the owner
constexpr int KIND_NEI = 1;   // internal face, this cell is the neighbour
constexpr int KIND_WALL = 2;  // Maxwell wall (fixedGradient gBarPvol, maxwellWall gSurf)
constexpr int KIND_CYC = 3;   // cyclic (coupled) patch face
constexpr int MAXS = 8;       // faces per cell handled by k_cells
constexpr int NAXMAX = 192;   // per-axis exp table entries (3 axes x 64 points)
constexpr int MAXNT = 512;     // workgroup size cap (256 VGPRs per lane available)
constexpr int MAXWAVES = MAXNT / 64;

// One (cell, face) incidence; a cell's slots are in ascending face order, the order in which the
// reference's face loops touch that cell.
struct alignas(16) Slot {
    double L[3];    // least-squares vector of this cell for this face
    double S[3];    // face area vector
    double dCs[3];  // Cf - C(this cell)
    double dCo[3];  // Cf - C(other cell) | wall: unit normal Sf/|Sf|
    double aux;     // wall: 1/deltaCoeffs | cyclic: patch weight w
    int32_t face, other, kind, bidx;
    double pad;
};
static_assert(sizeof(Slot) == 128, "slot size");

// Equilibrium constants of one face for the current step.
struct alignas(16) FaceEq {
    double U[3];
    double invRT;
    double pre;     // rho / (2 pi R T)^(D/2)
    double qc[3];   // (1 - Pr) q / (5 rho (R T)^2)
    double RT;
    double r;       // S7 relaxation factor h / (2 tau + h)
    double rho_w;   // wall density (walls only)
    double pad[5];
};
static_assert(sizeof(FaceEq) == 128, "faceeq size");

// Wall face constants: wall velocity/temperature, inComingByRho and the moments of the incoming
// Maxwellian (per unit rho_w) about the wall velocity.
struct alignas(16) WallC {
    double U[3];
    double T;
    double incoming;
    double pre;     // 1 / (2 pi R T_w)^(D/2)
    double A[14];   // rho, rhoU(D), rhoE, Q(D), P(D(D+1)/2) of the incoming part
    double pad[2];
};
static_assert(sizeof(WallC) == 176, "wallc size");

struct EqC {
    double U[3];
    double invRT, pre, qc[3], RT;
};

// One discrete velocity: xi, weight, |xi|^2 and the packed axis indices ix | iy << 8 | iz << 16.
struct alignas(16) VelRec {
    double x, y, z, w, xi2;
    int32_t idx, pad;
};
static_assert(sizeof(VelRec) == 48, "velrec size");

// Per-face data of k_fmom.
struct alignas(16) FaceInfo {
    double S[3];
    int32_t own, kind, bidx, pad;
};
static_assert(sizeof(FaceInfo) == 48, "faceinfo size");

template <typename TS>
struct P {
    int D, nc, nf, nvp, nxi, nax, K, nt, niter, nw;
    double R, Pr, omega, Tref, kc;  // kc = K + 3 - D
    const VelRec *__restrict__ vel;
    const double *__restrict__ axis;
    const int *__restrict__ cs;
    const Slot *__restrict__ slots;
    const double *__restrict__ cV;
    const int *__restrict__ ccase;
    const double *__restrict__ dtc;
    const double *__restrict__ muc;
    double *crho, *cU, *cT, *cq, *ctau;
    double *frho, *fU, *fT, *fq;
    FaceEq *feq;
    WallC *wall;
    double *wexp;
    double *qwall, *swall;
    const FaceInfo *__restrict__ finfo;
    TS *fg, *fh;        // face values [face][velocity] (S2/S4 result, before S7)
    const TS *gb_in, *hb_in;
    TS *gb_out, *hb_out;
    const TS *gBg_in, *gBh_in;
    TS *gBg_out, *gBh_out;
    int *flag;
};

#define DEVI __device__ __forceinline__

template <int D>
struct NMom {
    static constexpr int cons = D + 2;                 // rho, rhoU, rhoE
    static constexpr int v = D + 2 + D + D * (D + 1) / 2;  // + Q + P
};

// xi . S = (x Sx + y Sy) + z Sz without FMA contraction, as the reference computes it: on faces
// with |Sx| = |Sy| and velocities with |xi_x| = |xi_y| the two products cancel exactly to 0, which
// selects the averaging branch of the upwind test; a contracted product would leave the rounding
// error of one product and pick an upwind side instead.
DEVI double dotnc(double ax, double ay, double az, const double *b)
{
#pragma clang fp contract(off)
    return (ax * b[0] + ay * b[1]) + az * b[2];
}

template <typename TS>
DEVI double ld(const TS *a, size_t i)
{
    return (double)a[i];
}

struct Vel {
    double x, y, z, w, xi2;
    int ix, iy, iz;
};

template <typename TS>
DEVI Vel load_vel(const P<TS> &p, int k)
{
    const VelRec r = p.vel[k];
    Vel v;
    v.x = r.x;
    v.y = r.y;
    v.z = r.z;
    v.w = r.w;
    v.xi2 = r.xi2;
    v.ix = r.idx & 0xff;
    v.iy = (r.idx >> 8) & 0xff;
    v.iz = (r.idx >> 16) & 0xff;
    return v;
}

DEVI void eq_consts(double rho, const double *U, double T, const double *q, double R, double Pr, int D,
                    EqC &e)
{
    e.U[0] = U[0];
    e.U[1] = U[1];
    e.U[2] = U[2];
    e.RT = R * T;
    e.invRT = 1.0 / e.RT;
    double tp = 2.0 * PI * e.RT;
    e.pre = (D == 3) ? rho / (tp * sqrt(tp)) : rho / tp;
    double den = 5.0 * rho * R * T * R * T;
    double f = (1.0 - Pr) / den;
    e.qc[0] = f * q[0];
    e.qc[1] = f * q[1];
    e.qc[2] = f * q[2];
}

// Shakhov equilibrium (discreteVelocity.C:1016-1044) with the Gaussian from an axis table.
template <int D, bool H>
DEVI void eq_eval(const EqC &e, const double *tab, int nxi, const Vel &v, double kc, int K, double &geq,
                  double &heq)
{
    double cx = e.U[0] - v.x, cy = e.U[1] - v.y, cz = e.U[2] - v.z;
    double c2 = (cx * cx + cy * cy + cz * cz) * e.invRT;
    double pc = -(cx * e.qc[0] + cy * e.qc[1] + cz * e.qc[2]);
    double gB = e.pre * tab[v.ix] * tab[nxi + v.iy];
    if (D == 3) gB *= tab[2 * nxi + v.iz];
    geq = (1.0 + pc * ((c2 - D) - 2.0)) * gB;
    if (H) heq = (kc + pc * ((c2 - D) * kc - 2.0 * K)) * gB * e.RT;
    else heq = 0.0;
}

// Least-squares gradient of gbar (and hbar) at one cell for velocity k
// (leastSquaresGrad: owner += pVec (phi_N - phi_P), neighbour -= nVec (phi_N - phi_P),
// non-coupled boundary += pVec_b (phi_b - phi_P) with phi_b = phi_P + gradient_b / deltaCoeffs_b).
// The stencil values of up to MAXG faces are loaded before any is used, so their
// memory latencies overlap.
constexpr int MAXG = 6;

template <int D, bool H, typename TS>
DEVI void grad_cell(const P<TS> &p, int cell, int k, double pg, double ph, double *G, double *Hg)
{
#pragma unroll
    for (int d = 0; d < D; ++d) {
        G[d] = 0.0;
        Hg[d] = 0.0;
    }
    const int s0 = p.cs[cell], ns = p.cs[cell + 1] - s0;
    double vg[MAXG], vh[MAXG];
#pragma unroll
    for (int j = 0; j < MAXG; ++j) {
        vg[j] = 0.0;
        vh[j] = 0.0;
        if (j < ns) {
            const Slot &sl = p.slots[s0 + j];
            const bool wall = sl.kind == KIND_WALL;
            const size_t row = (size_t)(wall ? sl.bidx : sl.other) * p.nvp + k;
            vg[j] = ld(wall ? p.gBg_in : p.gb_in, row);
            if (H) vh[j] = ld(wall ? p.gBh_in : p.hb_in, row);
        }
    }
#pragma unroll
    for (int j = 0; j < MAXG; ++j) {
        if (j < ns) {
            const Slot &sl = p.slots[s0 + j];
            double dg, dh = 0.0;
            if (sl.kind == KIND_WALL) {
                dg = (pg + vg[j] * sl.aux) - pg;
                if (H) dh = (ph + vh[j] * sl.aux) - ph;
            } else {
                dg = vg[j] - pg;
                if (H) dh = vh[j] - ph;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                G[d] = fma(sl.L[d], dg, G[d]);
                if (H) Hg[d] = fma(sl.L[d], dh, Hg[d]);
            }
        }
    }
    for (int s = s0 + MAXG; s < s0 + ns; ++s) {
        const Slot &sl = p.slots[s];
        const bool wall = sl.kind == KIND_WALL;
        const size_t row = (size_t)(wall ? sl.bidx : sl.other) * p.nvp + k;
        const double a = ld(wall ? p.gBg_in : p.gb_in, row);
        const double b = H ? ld(wall ? p.gBh_in : p.hb_in, row) : 0.0;
        const double dg = wall ? (pg + a * sl.aux) - pg : a - pg;
        const double dh = wall ? (ph + b * sl.aux) - ph : b - ph;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            G[d] = fma(sl.L[d], dg, G[d]);
            if (H) Hg[d] = fma(sl.L[d], dh, Hg[d]);
        }
    }
}

template <int D>
DEVI double dotg(const double *G, const double *dC, const double *hx)
{
    double r = G[0] * (dC[0] - hx[0]) + G[1] * (dC[1] - hx[1]);
    if (D == 3) r += G[2] * (dC[2] - hx[2]);
    return r;
}

// Accumulate the face/cell moments of one velocity about reference velocity Uh.
template <int D, bool H>
DEVI void acc_moments(double *a, double g, double h, const Vel &v, const double *Uh)
{
    const double xi[3] = {v.x, v.y, v.z};
    double wg = v.w * g;
    a[0] += wg;
#pragma unroll
    for (int d = 0; d < D; ++d) a[1 + d] = fma(wg, xi[d], a[1 + d]);
    double hw = 0.5 * v.w;
    double e = g * v.xi2;
    if (H) e += h;
    a[D + 1] = fma(hw, e, a[D + 1]);
    double ch[3];
    double cc = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ch[d] = xi[d] - Uh[d];
        cc = fma(ch[d], ch[d], cc);
    }
    double A = cc * g;
    if (H) A += h;
#pragma unroll
    for (int d = 0; d < D; ++d) a[D + 2 + d] = fma(hw * ch[d], A, a[D + 2 + d]);
    double hwg = hw * g;
    int m = 2 * D + 2;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = i; j < D; ++j) {
            a[m] = fma(hwg * ch[i], ch[j], a[m]);
            ++m;
        }
}

// Macroscopic state from the reduced moments (fvDVM.C:493-522 / 694-727).
// Returns rho, U, T, tau, q (q before the bar-to-original correction factor).
template <int D>
DEVI void finish_moments(const double *a, const double *Uh, double R, int K, double mu, double omega,
                         double Tref, double &rho, double *U, double &T, double &tau, double *q)
{
    rho = a[0];
    double u2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) U[d] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        U[d] = a[1 + d] / rho;
        u2 += U[d] * U[d];
    }
    double rE = a[D + 1];
    T = (rE - 0.5 * rho * u2) / ((K + 3) / 2.0 * R * rho);
    tau = mu * exp(omega * log(T / Tref)) / rho / T / R;
    double dl[3] = {0, 0, 0}, B[3] = {0, 0, 0};
    double dd = 0.0, bd = 0.0, uhU = 0.0, uh2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dl[d] = U[d] - Uh[d];
        dd += dl[d] * dl[d];
        B[d] = 0.5 * (a[1 + d] - rho * Uh[d]);
        bd += B[d] * dl[d];
        uhU += Uh[d] * a[1 + d];
        uh2 += Uh[d] * Uh[d];
    }
    double A = rE - uhU + 0.5 * uh2 * rho;
    double M = 0.5 * rho;
    double Pm[3][3];
    {
        int m = 2 * D + 2;
        for (int i = 0; i < D; ++i)
            for (int j = i; j < D; ++j) {
                Pm[i][j] = a[m];
                Pm[j][i] = a[m];
                ++m;
            }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) q[i] = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double pd = 0.0;
        for (int j = 0; j < D; ++j) pd += Pm[i][j] * dl[j];
        q[i] = a[D + 2 + i] - 2.0 * pd + dd * B[i] - dl[i] * A + 2.0 * dl[i] * bd - dl[i] * dd * M;
    }
}

// Deterministic block sum of N per-thread values; result in out[0..N).
template <int N>
DEVI void block_sum(double *v, double *lds, double *out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double x = v[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[i] = x;
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds[wave * N + i] = v[i];
    }
    __syncthreads();
    if ((int)threadIdx.x < N) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += lds[w * N + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}


// ---------------------------------------------------------------------------------------------
// k_recon: S1 state, S2 gradient and upwind face reconstruction (discreteVelocity.C:412-691)
// One workgroup per cell, all velocities.  For every face and velocity the upwind cell writes the
// face value: the owner for xi.S >= VSMALL, the neighbour for xi.S < -VSMALL; the owner writes the
// average branch (|xi.S| < VSMALL) and the cyclic incoming value, which need the other cell's
// gradient.  Wall faces get their outgoing values; incoming wall values come from rho_w later.
// Also writes the next step's fixedGradient gradient n . grad at wall faces.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_recon(P<TS> p)
{
    const int c = blockIdx.x;
    const double dt = p.dtc[p.ccase[c]];
    const int s0 = p.cs[c], ns = p.cs[c + 1] - s0;
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel(p, k);
        const double hx[3] = {0.5 * v.x * dt, 0.5 * v.y * dt, 0.5 * v.z * dt};
        const size_t ck = (size_t)c * p.nvp + k;
        const double gc = ld(p.gb_in, ck);
        const double hc = H ? ld(p.hb_in, ck) : 0.0;
        double Gc[3] = {0, 0, 0}, Hc[3] = {0, 0, 0};
        grad_cell<D, H>(p, c, k, gc, hc, Gc, Hc);
        for (int j = 0; j < ns; ++j) {
            const Slot &S = p.slots[s0 + j];
            const double s = dotnc(v.x, v.y, v.z, S.S);
            const size_t fk = (size_t)S.face * p.nvp + k;
            if (S.kind == KIND_WALL) {
                const size_t bi = (size_t)S.bidx * p.nvp + k;
                double ng = Gc[0] * S.dCo[0] + Gc[1] * S.dCo[1];
                if (D == 3) ng += Gc[2] * S.dCo[2];
                p.gBg_out[bi] = (TS)ng;
                if (H) {
                    double nh = Hc[0] * S.dCo[0] + Hc[1] * S.dCo[1];
                    if (D == 3) nh += Hc[2] * S.dCo[2];
                    p.gBh_out[bi] = (TS)nh;
                }
                if (s > 0) {
                    p.fg[fk] = (TS)(gc + dotg<D>(Gc, S.dCs, hx));
                    if (H) p.fh[fk] = (TS)(hc + dotg<D>(Hc, S.dCs, hx));
                }
            } else if (S.kind == KIND_NEI) {
                if (s < -VSMALL) {
                    p.fg[fk] = (TS)(gc + dotg<D>(Gc, S.dCs, hx));
                    if (H) p.fh[fk] = (TS)(hc + dotg<D>(Hc, S.dCs, hx));
                }
            } else {
                const bool cyc = S.kind == KIND_CYC;
                const bool up = cyc ? (s > VSMALL) : (s >= VSMALL);
                const bool need = !up && (cyc || !(s < -VSMALL));
                const double dcg = dotg<D>(Gc, S.dCs, hx);
                const double dch = H ? dotg<D>(Hc, S.dCs, hx) : 0.0;
                double vg = gc + dcg, vh = hc + dch;
                if (__any(need)) {
                    const size_t ok = (size_t)S.other * p.nvp + k;
                    double go = ld(p.gb_in, ok);
                    double ho = H ? ld(p.hb_in, ok) : 0.0;
                    double Go[3] = {0, 0, 0}, Ho[3] = {0, 0, 0};
                    grad_cell<D, H>(p, S.other, k, go, ho, Go, Ho);
                    double dog, doh = 0.0;
                    if (cyc) {
                        // coupled boundary values w phi_P + (1-w) phi_N, same for the gradient,
                        // at distance Cf - Cf - xi dt/2 (discreteVelocity.C:645-667)
                        const double wb = S.aux;
                        double Gb[3] = {0, 0, 0}, Hb[3] = {0, 0, 0};
                        for (int d = 0; d < D; ++d) {
                            Gb[d] = wb * Gc[d] + (1.0 - wb) * Go[d];
                            Hb[d] = wb * Hc[d] + (1.0 - wb) * Ho[d];
                        }
                        const double z3[3] = {0, 0, 0};
                        go = wb * gc + (1.0 - wb) * go;
                        dog = dotg<D>(Gb, z3, hx);
                        if (H) {
                            ho = wb * hc + (1.0 - wb) * ho;
                            doh = dotg<D>(Hb, z3, hx);
                        }
                    } else {
                        dog = dotg<D>(Go, S.dCo, hx);
                        if (H) doh = dotg<D>(Ho, S.dCo, hx);
                    }
                    if (need) {
                        if (cyc && s < -VSMALL) {
                            vg = go + dog;
                            vh = ho + doh;
                        } else if (cyc) {
                            // 0.5*(phi_P + grad_P.d_P + phi_b + grad_b.d_b)
                            vg = 0.5 * (((gc + dcg) + go) + dog);
                            vh = 0.5 * (((hc + dch) + ho) + doh);
                        } else {
                            // 0.5*(phi_N + grad_N.d_N + phi_P + grad_P.d_P) (discreteVelocity.C:515-528)
                            vg = 0.5 * (((go + dog) + gc) + dcg);
                            vh = 0.5 * (((ho + doh) + hc) + dch);
                        }
                    }
                }
                if (up || need) {
                    p.fg[fk] = (TS)vg;
                    if (H) p.fh[fk] = (TS)vh;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// k_fmom: S3 wall density, S4 incoming wall Maxwellian, S6 face moments (fvDVM.C:347-582) and
// S7 face relaxation (discreteVelocity.C:819-932).  One workgroup per face, all velocities:
// the first pass reduces the moments of the stored face values, the second rewrites them
// relaxed towards the face equilibrium (incoming wall values: rho_w times the wall Maxwellian,
// unrelaxed).  With NIT > 0 a thread keeps its face values in registers between the passes.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, typename TS, int NIT>
__global__ __launch_bounds__(MAXNT) void k_fmom(P<TS> p)
{
    constexpr int NM = NMom<D>::v + 1;  // + wall outGoing
    constexpr int NR = NIT > 0 ? NIT : 1;
    constexpr int NL = NIT > 0 ? NIT : 64;
    __shared__ double red[MAXWAVES * NM];
    __shared__ double tot[NM];
    __shared__ double tabF[NAXMAX];
    __shared__ double tabW[NAXMAX];
    __shared__ double fin[16];
    const int f = blockIdx.x;
    const FaceInfo fi = p.finfo[f];
    const int cas = p.ccase[fi.own];
    const double dt = p.dtc[cas];
    const bool wall = fi.kind == KIND_WALL;
    const TS *__restrict__ fg = p.fg;
    const TS *__restrict__ fh = p.fh;
    double Uh[3];
    if (wall) {
        const WallC &W = p.wall[fi.bidx];
        Uh[0] = W.U[0]; Uh[1] = W.U[1]; Uh[2] = W.U[2];
        for (int t = threadIdx.x; t < p.nax; t += blockDim.x) tabW[t] = p.wexp[(size_t)fi.bidx * p.nax + t];
    } else {
        Uh[0] = p.fU[3 * f]; Uh[1] = p.fU[3 * f + 1]; Uh[2] = p.fU[3 * f + 2];
    }
    double acc[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) acc[m] = 0.0;
    double vg[NR], vh[NR];
#pragma unroll(NIT > 0 ? NIT : 2)
    for (int it = 0; it < NL; ++it) {
        if (it < p.niter) {
            const int k = threadIdx.x + it * p.nt;
            const Vel v = load_vel(p, k);
            const double s = dotnc(v.x, v.y, v.z, fi.S);
            const bool in = wall && !(s > 0);   // incoming wall values: rho_w times the wall Maxwellian
            const size_t fk = (size_t)f * p.nvp + k;
            const double gf = in ? 0.0 : ld(fg, fk);
            const double hf = (H && !in) ? ld(fh, fk) : 0.0;
            if (NIT > 0) {
                vg[NIT > 0 ? it : 0] = gf;
                vh[NIT > 0 ? it : 0] = hf;
            }
            if (!in) {
                acc_moments<D, H>(acc, gf, hf, v, Uh);
                if (wall) acc[NM - 1] = fma(v.w * s, gf, acc[NM - 1]);
            }
        }
    }
    block_sum<NM>(acc, red, tot);
    if (threadIdx.x == 0) {
        double a[NM];
        for (int m = 0; m < NM; ++m) a[m] = tot[m];
        double rho_w = 0.0;
        if (wall) {
            const WallC &W = p.wall[fi.bidx];
            // S3: rho_w = outGoing / |inComingByRho| (calculatedMaxwellFvPatchField.C:139-164)
            rho_w = a[NM - 1] / fabs(W.incoming);
            // S4: incoming DFs are rho_w times the wall Maxwellian; add their moments
            for (int m = 0; m < NM - 1; ++m) a[m] = fma(rho_w, W.A[m], a[m]);
        }
        double rho, U[3], T, tau, q[3];
        const double mu = p.muc[cas];
        finish_moments<D>(a, Uh, p.R, p.K, mu, p.omega, p.Tref, rho, U, T, tau, q);
        const double t2 = 2.0 * tau;
        const double fq = t2 / (t2 + 0.5 * dt * p.Pr);
        for (int d = 0; d < 3; ++d) q[d] = fq * q[d];
        p.frho[f] = rho;
        p.fT[f] = T;
        for (int d = 0; d < 3; ++d) {
            p.fU[3 * f + d] = U[d];
            p.fq[3 * f + d] = q[d];
        }
        EqC e;
        eq_consts(rho, U, T, q, p.R, p.Pr, D, e);
        const double hh = 0.5 * dt;
        FaceEq fe;
        for (int d = 0; d < 3; ++d) {
            fe.U[d] = e.U[d];
            fe.qc[d] = e.qc[d];
        }
        fe.invRT = e.invRT;
        fe.pre = e.pre;
        fe.RT = e.RT;
        fe.r = hh / (2 * tau + hh);
        fe.rho_w = rho_w;
        for (int d = 0; d < 5; ++d) fe.pad[d] = 0.0;
        p.feq[f] = fe;
        if (wall) {
            // qWall, stressWall (fvDVM.C:539-581): moments about the wall velocity
            const WallC &W = p.wall[fi.bidx];
            const double fw = t2 / (t2 + 0.5 * dt * p.Pr);
            const double fs = t2 / (t2 + 0.5 * dt);
            double Pm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            int m = 2 * D + 2;
            for (int i = 0; i < D; ++i)
                for (int jj = i; jj < D; ++jj) {
                    Pm[i][jj] = a[m];
                    Pm[jj][i] = a[m];
                    ++m;
                }
            double rU[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) rU[d] = a[1 + d];
            for (int d = 0; d < 3; ++d) p.qwall[3 * fi.bidx + d] = d < D ? fw * a[D + 2 + d] : 0.0;
            for (int i = 0; i < 3; ++i)
                for (int jj = 0; jj < 3; ++jj) {
                    double sv = 2.0 * Pm[i][jj] + W.U[i] * rU[jj] + W.U[jj] * rU[i] - W.U[i] * W.U[jj] * rho;
                    p.swall[9 * fi.bidx + 3 * i + jj] = fs * sv;
                }
        }
        fin[0] = e.U[0]; fin[1] = e.U[1]; fin[2] = e.U[2];
        fin[3] = e.invRT; fin[4] = e.pre; fin[5] = e.qc[0]; fin[6] = e.qc[1]; fin[7] = e.qc[2];
        fin[8] = e.RT; fin[9] = fe.r; fin[10] = rho_w;
    }
    __syncthreads();
    // per-axis exp table of the face equilibrium
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = fin[ax] - p.axis[t - ax * p.nxi];
        tabF[t] = exp(-(u * u) * (0.5 * fin[3]));
    }
    __syncthreads();
    EqC e;
    e.U[0] = fin[0]; e.U[1] = fin[1]; e.U[2] = fin[2];
    e.invRT = fin[3]; e.pre = fin[4]; e.qc[0] = fin[5]; e.qc[1] = fin[6]; e.qc[2] = fin[7]; e.RT = fin[8];
    const double r = fin[9], rho_w = fin[10];
    double Tw = 1.0, prew = 0.0;
    if (wall) {
        Tw = p.wall[fi.bidx].T;
        prew = p.wall[fi.bidx].pre;
    }
    TS *__restrict__ fgo = p.fg;
    TS *__restrict__ fho = p.fh;
#pragma unroll(NIT > 0 ? NIT : 2)
    for (int it = 0; it < NL; ++it) {
        if (it < p.niter) {
            const int k = threadIdx.x + it * p.nt;
            const Vel v = load_vel(p, k);
            const double s = dotnc(v.x, v.y, v.z, fi.S);
            const size_t fk = (size_t)f * p.nvp + k;
            double gf, hf = 0.0;
            if (wall && !(s > 0)) {
                // S4 incoming Maxwellian, left unrelaxed by S7 (maxwellWall operators are no-ops)
                double m = prew * tabW[v.ix] * tabW[p.nxi + v.iy];
                if (D == 3) m *= tabW[2 * p.nxi + v.iz];
                gf = rho_w * m;
                if (H) hf = gf * (p.R * Tw) * p.kc;
            } else {
                gf = NIT > 0 ? vg[NIT > 0 ? it : 0] : ld(fgo, fk);
                if (H) hf = NIT > 0 ? vh[NIT > 0 ? it : 0] : ld(fho, fk);
                double gq, hq;
                eq_eval<D, H>(e, tabF, p.nxi, v, p.kc, p.K, gq, hq);
                gf = (1.0 - r) * gf + r * gq;
                if (H) hf = (1.0 - r) * hf + r * hq;
            }
            fgo[fk] = (TS)gf;
            if (H) fho[fk] = (TS)hf;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// k_cells: S8 g~ update with the relaxed face values, S9 cell moments, then S1 of the next step
// One workgroup per cell, all velocities.  With NIT > 0 each thread keeps its NIT new g~ values
// in registers until the cell moments are known; with NIT = 0 it writes them to its own slots of
// the output row and rewrites them there as gbar.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, typename TS, int NIT>
__global__ __launch_bounds__(MAXNT) void k_cells(P<TS> p)
{
    constexpr int NR = NIT > 0 ? NIT : 1;
    constexpr int NL = NIT > 0 ? NIT : 64;
    constexpr int NC = NMom<D>::v;
    __shared__ double tabC[NAXMAX];
    __shared__ double tabN[NAXMAX];
    __shared__ double red[MAXWAVES * NC];
    __shared__ double tot[NC];
    __shared__ double newm[16];

    const int c = blockIdx.x;
    const int cas = p.ccase[c];
    const double dt = p.dtc[cas];
    const double V = p.cV[c];
    const int s0 = p.cs[c];
    const int ns = p.cs[c + 1] - s0;
    const TS *__restrict__ fg = p.fg;
    const TS *__restrict__ fh = p.fh;
    const TS *__restrict__ gbi = p.gb_in;
    const TS *__restrict__ hbi = p.hb_in;

    // state that formed gbar: macros of the previous step
    const double Uc[3] = {p.cU[3 * c], p.cU[3 * c + 1], p.cU[3 * c + 2]};
    const double qcv[3] = {p.cq[3 * c], p.cq[3 * c + 1], p.cq[3 * c + 2]};
    EqC ec;
    eq_consts(p.crho[c], Uc, p.cT[c], qcv, p.R, p.Pr, D, ec);
    const double rc = 1.5 * dt / (2.0 * p.ctau[c] + dt);
    const double inv1mr = 1.0 / (1.0 - rc);
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = Uc[ax] - p.axis[t - ax * p.nxi];
        tabC[t] = exp(-(u * u) * (0.5 * ec.invRT));
    }
    __syncthreads();

    double gnew[NR], hnew[NR];
    double acc[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] = 0.0;

#pragma unroll(NIT > 0 ? NIT : 2)
    for (int it = 0; it < NL; ++it) {
        if (it < p.niter) {
            const int k = threadIdx.x + it * p.nt;
            const Vel v = load_vel(p, k);
            const size_t ck = (size_t)c * p.nvp + k;
            const double gb = ld(gbi, ck);
            const double hb = H ? ld(hbi, ck) : 0.0;
            // relaxed face values of every face of the cell, loaded together
            double fgv[MAXS], fhv[MAXS];
#pragma unroll
            for (int j = 0; j < MAXS; ++j) {
                fgv[j] = 0.0;
                fhv[j] = 0.0;
                if (j < ns) {
                    const size_t fk = (size_t)p.slots[s0 + j].face * p.nvp + k;
                    fgv[j] = ld(fg, fk);
                    if (H) fhv[j] = ld(fh, fk);
                }
            }
            // g~ of this step from gbar (inverse of S1)
            double geq, heq;
            eq_eval<D, H>(ec, tabC, p.nxi, v, p.kc, p.K, geq, heq);
            const double gt = (gb - rc * geq) * inv1mr;
            const double ht = H ? (hb - rc * heq) * inv1mr : 0.0;
            // S8 (discreteVelocity.C:937-975): -1/3 g~ + 4/3 gbar, then the faces in face order
            double ag = -1.0 / 3 * gt + 4.0 / 3 * gb;
            double ah = H ? (-1.0 / 3 * ht + 4.0 / 3 * hb) : 0.0;
#pragma unroll
            for (int j = 0; j < MAXS; ++j) {
                if (j < ns) {
                    const Slot &S = p.slots[s0 + j];
                    const double s = dotnc(v.x, v.y, v.z, S.S);
                    const double tg = s * fgv[j] * dt / V;
                    if (S.kind == KIND_NEI) ag += tg;
                    else ag -= tg;
                    if (H) {
                        const double th = s * fhv[j] * dt / V;
                        if (S.kind == KIND_NEI) ah += th;
                        else ah -= th;
                    }
                }
            }
            if (NIT > 0) {
                gnew[NIT > 0 ? it : 0] = ag;
                hnew[NIT > 0 ? it : 0] = ah;
            } else {
                p.gb_out[ck] = (TS)ag;
                if (H) p.hb_out[ck] = (TS)ah;
            }
            acc_moments<D, H>(acc, ag, ah, v, Uc);
        }
    }

    block_sum<NC>(acc, red, tot);
    if (threadIdx.x == 0) {
        double rho, U[3], T, tau, q[3];
        const double mu = p.muc[cas];
        finish_moments<D>(tot, Uc, p.R, p.K, mu, p.omega, p.Tref, rho, U, T, tau, q);
        const double fqc = 2.0 * tau / (2.0 * tau + dt * p.Pr);
        for (int d = 0; d < 3; ++d) q[d] = fqc * q[d];
        p.crho[c] = rho;
        p.cT[c] = T;
        p.ctau[c] = tau;
        for (int d = 0; d < 3; ++d) {
            p.cU[3 * c + d] = U[d];
            p.cq[3 * c + d] = q[d];
        }
        if (!(rho > 0.0) || !(T > 0.0) || !isfinite(rho) || !isfinite(T) || !isfinite(tau)) p.flag[0] = 1;
        newm[0] = rho; newm[1] = U[0]; newm[2] = U[1]; newm[3] = U[2]; newm[4] = T;
        newm[5] = q[0]; newm[6] = q[1]; newm[7] = q[2]; newm[8] = tau;
    }
    __syncthreads();
    EqC en;
    {
        const double Un[3] = {newm[1], newm[2], newm[3]};
        const double qn[3] = {newm[5], newm[6], newm[7]};
        eq_consts(newm[0], Un, newm[4], qn, p.R, p.Pr, D, en);
        for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
            const int ax = t / p.nxi;
            const double u = Un[ax] - p.axis[t - ax * p.nxi];
            tabN[t] = exp(-(u * u) * (0.5 * en.invRT));
        }
    }
    __syncthreads();
    // S1 of the next step (discreteVelocity.C:393-406)
    const double rn = 1.5 * dt / (2.0 * newm[8] + dt);
#pragma unroll(NIT > 0 ? NIT : 2)
    for (int it = 0; it < NL; ++it) {
        if (it < p.niter) {
            const int k = threadIdx.x + it * p.nt;
            const Vel v = load_vel(p, k);
            double geq, heq;
            eq_eval<D, H>(en, tabN, p.nxi, v, p.kc, p.K, geq, heq);
            const size_t ck = (size_t)c * p.nvp + k;
            const double gn = NIT > 0 ? gnew[NIT > 0 ? it : 0] : ld(p.gb_out, ck);
            const double hn = !H ? 0.0 : (NIT > 0 ? hnew[NIT > 0 ? it : 0] : ld(p.hb_out, ck));
            p.gb_out[ck] = (TS)((1.0 - rn) * gn + rn * geq);
            if (H) p.hb_out[ck] = (TS)((1.0 - rn) * hn + rn * heq);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// auxiliary kernels
// ---------------------------------------------------------------------------------------------

// initDFtoEq (discreteVelocity.C:220-249): g~ = Shakhov(rho, U, T, q = 0), written raw.
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_init_raw(P<TS> p)
{
    __shared__ double tab[NAXMAX];
    const int c = blockIdx.x;
    const double U[3] = {p.cU[3 * c], p.cU[3 * c + 1], p.cU[3 * c + 2]};
    const double q0[3] = {0, 0, 0};
    EqC e;
    eq_consts(p.crho[c], U, p.cT[c], q0, p.R, p.Pr, D, e);
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = U[ax] - p.axis[t - ax * p.nxi];
        tab[t] = exp(-(u * u) * (0.5 * e.invRT));
    }
    __syncthreads();
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel(p, k);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        p.gb_out[ck] = (TS)geq;
        if (H) p.hb_out[ck] = (TS)heq;
    }
}

// Converts between g~ and gbar = (1-r) g~ + r geq of the cell macros.
// mode 0: g~ -> gbar with dt_new; mode 1: gbar(dt_old) -> gbar(dt_new); mode 2: gbar(dt_old) -> g~.
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_convert(P<TS> p, const double *dt_old, const double *dt_new, int mode)
{
    __shared__ double tab[NAXMAX];
    const int c = blockIdx.x;
    const int cas = p.ccase[c];
    const double U[3] = {p.cU[3 * c], p.cU[3 * c + 1], p.cU[3 * c + 2]};
    const double q[3] = {p.cq[3 * c], p.cq[3 * c + 1], p.cq[3 * c + 2]};
    EqC e;
    eq_consts(p.crho[c], U, p.cT[c], q, p.R, p.Pr, D, e);
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = U[ax] - p.axis[t - ax * p.nxi];
        tab[t] = exp(-(u * u) * (0.5 * e.invRT));
    }
    __syncthreads();
    const double tau = p.ctau[c];
    double ro = 0.0, rn = 0.0;
    if (mode != 0) ro = 1.5 * dt_old[cas] / (2.0 * tau + dt_old[cas]);
    if (mode != 2) rn = 1.5 * dt_new[cas] / (2.0 * tau + dt_new[cas]);
    const double inv1mr = 1.0 / (1.0 - ro);
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel(p, k);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        double g = ld(p.gb_in, ck), h = H ? ld(p.hb_in, ck) : 0.0;
        if (mode != 0) {
            g = (g - ro * geq) * inv1mr;
            h = (h - ro * heq) * inv1mr;
        }
        if (mode != 2) {
            g = (1.0 - rn) * g + rn * geq;
            h = (1.0 - rn) * h + rn * heq;
        }
        p.gb_out[ck] = (TS)g;
        if (H) p.hb_out[ck] = (TS)h;
    }
}

// Wall set-up: axis tables of the wall Maxwellian, inComingByRho (fvDVM.C:263-309) and the
// moments of the incoming part per unit rho_w.
template <int D, typename TS>
__global__ __launch_bounds__(256) void k_wall_init(P<TS> p, const int *wall_face)
{
    constexpr int NM = NMom<D>::v;
    __shared__ double tab[NAXMAX];
    __shared__ double red[MAXWAVES * (NM + 1)];
    __shared__ double tot[NM + 1];
    const int b = blockIdx.x;
    WallC &W = p.wall[b];
    const double *Sf = p.slots[wall_face[b]].S;
    const double RT = p.R * W.T;
    const double invRT = 1.0 / RT;
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = W.U[ax] - p.axis[t - ax * p.nxi];
        const double e = exp(-(u * u) * (0.5 * invRT));
        tab[t] = e;
        p.wexp[(size_t)b * p.nax + t] = e;
    }
    __syncthreads();
    double acc[NM + 1];
    for (int m = 0; m <= NM; ++m) acc[m] = 0.0;
    const double Uw[3] = {W.U[0], W.U[1], W.U[2]};
    for (int k = threadIdx.x; k < p.nvp; k += blockDim.x) {
        const Vel v = load_vel(p, k);
        const double s = dotnc(v.x, v.y, v.z, Sf);
        double m = W.pre * tab[v.ix] * tab[p.nxi + v.iy];
        if (D == 3) m *= tab[2 * p.nxi + v.iz];
        if (s < 0) acc[NM] = fma(-v.w * s, m, acc[NM]);
        if (s <= 0) acc_moments<D, true>(acc, m, m * RT * p.kc, v, Uw);
    }
    block_sum<NM + 1>(acc, red, tot);
    if (threadIdx.x == 0) {
        W.incoming = tot[NM];
        for (int m = 0; m < 14; ++m) W.A[m] = m < NM ? tot[m] : 0.0;
    }
}

// g~ of selected cells for get_df / save: mode-2 conversion into a scratch row buffer.
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_extract(P<TS> p, const int *cells, const double *dt_old, int raw,
                                                  double *g_out, double *h_out)
{
    __shared__ double tab[NAXMAX];
    const int c = cells[blockIdx.x];
    const int cas = p.ccase[c];
    const double U[3] = {p.cU[3 * c], p.cU[3 * c + 1], p.cU[3 * c + 2]};
    const double q[3] = {p.cq[3 * c], p.cq[3 * c + 1], p.cq[3 * c + 2]};
    EqC e;
    eq_consts(p.crho[c], U, p.cT[c], q, p.R, p.Pr, D, e);
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = U[ax] - p.axis[t - ax * p.nxi];
        tab[t] = exp(-(u * u) * (0.5 * e.invRT));
    }
    __syncthreads();
    const double ro = raw ? 0.0 : 1.5 * dt_old[cas] / (2.0 * p.ctau[c] + dt_old[cas]);
    const double inv1mr = 1.0 / (1.0 - ro);
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel(p, k);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        double g = ld(p.gb_in, ck), h = H ? ld(p.hb_in, ck) : 0.0;
        if (!raw) {
            g = (g - ro * geq) * inv1mr;
            h = (h - ro * heq) * inv1mr;
        }
        g_out[(size_t)blockIdx.x * p.nvp + k] = g;
        h_out[(size_t)blockIdx.x * p.nvp + k] = H ? h : 0.0;
    }
}

}  // namespace

// =============================================================================================
// host side
// =============================================================================================

struct dugks_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 2, nc = 0, nfi = 0, nf = 0, nb = 0, nw = 0, nv = 0, nvp = 0, nxi = 0, nax = 0, K = 0;
    int storage = DUGKS_STORAGE_FP64, n_cases = 1;
    bool has_h = true;
    int nt = 64, niter = 1, nit_max = 1, fmom_nit = 0;
    double R = 0, omega = 0, Tref = 0, mu_ref = 0, Pr = 0, xi_max = 0;
    std::vector<int> perm;          // device velocity -> reference velocity, -1 for padding
    std::vector<double> dcoef_int;  // internal deltaCoeffs (Courant number)
    std::vector<double> face_w;     // surfaceInterpolation weights
    std::vector<double> dt_host;    // staging of the per-case dt
    std::vector<int> patch_type, patch_start, patch_size, patch_wall0;
    std::vector<int> wall_face_slot;  // per wall: its slot index
    std::vector<int> ncell_slots;

    // device buffers
    VelRec *vel = nullptr;
    double *axis = nullptr;
    int *cs = nullptr, *ccase = nullptr, *flag = nullptr, *wfslot = nullptr;
    Slot *slots = nullptr;
    double *cV = nullptr, *dtc = nullptr, *dtp = nullptr, *muc = nullptr;
    double *crho = nullptr, *cU = nullptr, *cT = nullptr, *cq = nullptr, *ctau = nullptr;
    double *frho = nullptr, *fU = nullptr, *fT = nullptr, *fq = nullptr;
    FaceEq *feq = nullptr;
    WallC *wall = nullptr;
    double *wexp = nullptr, *qwall = nullptr, *swall = nullptr;
    FaceInfo *finfo = nullptr;
    void *fg = nullptr, *fh = nullptr;
    void *gb[2] = {nullptr, nullptr}, *hb[2] = {nullptr, nullptr};
    void *gB[2] = {nullptr, nullptr}, *hB[2] = {nullptr, nullptr};
    int cur = 0;         // buffer holding the current state
    int gcur = 0;        // boundary-gradient buffer to read
    int raw = 1;         // state holds g~ (1) or gbar prepared with dt_prep (0)
    int initialised = 0;
    std::vector<double> dt_prep;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int profiling = 0;
    std::vector<hipEvent_t> pev;     // 4 events per profiled step
    double prof_recon = 0.0, prof_fmom = 0.0, prof_cells = 0.0;
    int64_t prof_steps = 0;
    std::string err;
};

namespace {

#define HIPCHK(h, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                       \
            return DUGKS_ERR_DEVICE;                                                           \
        }                                                                                      \
    } while (0)

int fail(dugks_engine *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    return code;
}

template <typename T>
int dalloc(dugks_engine *h, T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc(") + std::to_string(sizeof(T) * n) + "): " + hipGetErrorString(e);
        *p = nullptr;
        return DUGKS_ERR_MEMORY;
    }
    return 0;
}

int dallocb(dugks_engine *h, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        *p = nullptr;
        return DUGKS_ERR_MEMORY;
    }
    return 0;
}

template <typename T>
int upload(dugks_engine *h, T *d, const T *src, size_t n)
{
    HIPCHK(h, hipMemcpy(d, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return 0;
}

size_t ts_size(const dugks_engine *h) { return h->storage == DUGKS_STORAGE_FP32 ? 4 : 8; }

template <typename TS>
P<TS> params(dugks_engine *h)
{
    P<TS> p;
    p.D = h->D; p.nc = h->nc; p.nf = h->nf; p.nvp = h->nvp; p.nxi = h->nxi; p.nax = h->nax; p.K = h->K;
    p.nt = h->nt; p.niter = h->niter; p.nw = h->nw;
    p.R = h->R; p.Pr = h->Pr; p.omega = h->omega; p.Tref = h->Tref; p.kc = (double)(h->K + 3 - h->D);
    p.vel = h->vel;
    p.axis = h->axis; p.cs = h->cs; p.slots = h->slots; p.cV = h->cV; p.ccase = h->ccase;
    p.dtc = h->dtc; p.muc = h->muc;
    p.crho = h->crho; p.cU = h->cU; p.cT = h->cT; p.cq = h->cq; p.ctau = h->ctau;
    p.frho = h->frho; p.fU = h->fU; p.fT = h->fT; p.fq = h->fq; p.feq = h->feq;
    p.wall = h->wall; p.wexp = h->wexp; p.qwall = h->qwall; p.swall = h->swall; p.finfo = h->finfo;
    p.fg = (TS *)h->fg; p.fh = (TS *)h->fh;
    p.gb_in = (const TS *)h->gb[h->cur]; p.hb_in = (const TS *)h->hb[h->cur];
    p.gb_out = (TS *)h->gb[1 - h->cur]; p.hb_out = (TS *)h->hb[1 - h->cur];
    p.gBg_in = (const TS *)h->gB[h->gcur]; p.gBh_in = (const TS *)h->hB[h->gcur];
    p.gBg_out = (TS *)h->gB[1 - h->gcur]; p.gBh_out = (TS *)h->hB[1 - h->gcur];
    p.flag = h->flag;
    return p;
}

// ---- launch dispatch over (D, H, TS, NIT) ----

template <int D, bool H, typename TS, int NIT>
void launch_cells_t(dugks_engine *h, const P<TS> &p)
{
    hipLaunchKernelGGL((k_cells<D, H, TS, NIT>), dim3(h->nc), dim3(h->nt), 0, h->stream, p);
}

template <int D, bool H, typename TS>
void launch_cells_n(dugks_engine *h, const P<TS> &p)
{
    switch (h->nit_max) {
    case 1: launch_cells_t<D, H, TS, 1>(h, p); break;
    case 2: launch_cells_t<D, H, TS, 2>(h, p); break;
    case 4: launch_cells_t<D, H, TS, 4>(h, p); break;
    case 8: launch_cells_t<D, H, TS, 8>(h, p); break;
    default: launch_cells_t<D, H, TS, 0>(h, p); break;
    }
}

template <int D, bool H, typename TS, int NIT>
void launch_fmom_t(dugks_engine *h, const P<TS> &p)
{
    hipLaunchKernelGGL((k_fmom<D, H, TS, NIT>), dim3(h->nf), dim3(h->nt), 0, h->stream, p);
}

template <int D, bool H, typename TS>
void launch_fmom_n(dugks_engine *h, const P<TS> &p)
{
    switch (h->fmom_nit) {
    case 1: launch_fmom_t<D, H, TS, 1>(h, p); break;
    case 2: launch_fmom_t<D, H, TS, 2>(h, p); break;
    case 4: launch_fmom_t<D, H, TS, 4>(h, p); break;
    case 8: launch_fmom_t<D, H, TS, 8>(h, p); break;
    case 16: launch_fmom_t<D, H, TS, 16>(h, p); break;
    case 32: launch_fmom_t<D, H, TS, 32>(h, p); break;
    default: launch_fmom_t<D, H, TS, 0>(h, p); break;
    }
}

template <int D, bool H, typename TS>
void launch_step_t(dugks_engine *h, hipEvent_t *ev)
{
    P<TS> p = params<TS>(h);
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    hipLaunchKernelGGL((k_recon<D, H, TS>), dim3(h->nc), dim3(h->nt), 0, h->stream, p);
    if (ev) (void)hipEventRecord(ev[1], h->stream);
    launch_fmom_n<D, H, TS>(h, p);
    if (ev) (void)hipEventRecord(ev[2], h->stream);
    launch_cells_n<D, H, TS>(h, p);
    if (ev) (void)hipEventRecord(ev[3], h->stream);
}

template <typename F>
void dispatch(dugks_engine *h, F &&f)
{
    const bool fp32 = h->storage == DUGKS_STORAGE_FP32;
    if (h->D == 3) {
        if (h->has_h) {
            if (fp32) f.template operator()<3, true, float>(); else f.template operator()<3, true, double>();
        } else {
            if (fp32) f.template operator()<3, false, float>(); else f.template operator()<3, false, double>();
        }
    } else {
        // in 2-D, K + 3 - D >= 1: h is always carried
        if (fp32) f.template operator()<2, true, float>(); else f.template operator()<2, true, double>();
    }
}

struct StepOp {
    dugks_engine *h;
    hipEvent_t *ev;
    template <int D, bool H, typename TS>
    void operator()() { launch_step_t<D, H, TS>(h, ev); }
};

struct InitOp {
    dugks_engine *h;
    template <int D, bool H, typename TS>
    void operator()()
    {
        P<TS> p = params<TS>(h);
        p.gb_out = (TS *)h->gb[h->cur];
        p.hb_out = (TS *)h->hb[h->cur];
        hipLaunchKernelGGL((k_init_raw<D, H, TS>), dim3(h->nc), dim3(h->nt), 0, h->stream, p);
        if (h->nw > 0)
            hipLaunchKernelGGL((k_wall_init<D, TS>), dim3(h->nw), dim3(256), 0, h->stream, p,
                               (const int *)h->wfslot);
    }
};

struct ConvertOp {
    dugks_engine *h;
    int mode;
    template <int D, bool H, typename TS>
    void operator()()
    {
        P<TS> p = params<TS>(h);
        hipLaunchKernelGGL((k_convert<D, H, TS>), dim3(h->nc), dim3(h->nt), 0, h->stream, p,
                           (const double *)h->dtp, (const double *)h->dtc, mode);
    }
};

struct ExtractOp {
    dugks_engine *h;
    const int *cells;
    int n;
    double *g, *hh;
    template <int D, bool H, typename TS>
    void operator()()
    {
        P<TS> p = params<TS>(h);
        hipLaunchKernelGGL((k_extract<D, H, TS>), dim3(n), dim3(h->nt), 0, h->stream, p, cells,
                           (const double *)h->dtp, h->raw, g, hh);
    }
};

void choose_launch(int nv, int &nt, int &niter, int &nvp)
{
    int best_pad = 1 << 30, best_it = 0, best_nt = 0;
    for (int it = 1; it <= 64; ++it) {
        int t = (nv + it - 1) / it;
        t = (t + 63) / 64 * 64;
        if (t > MAXNT) continue;
        int pad = t * it - nv;
        if (pad < best_pad) {
            best_pad = pad;
            best_it = it;
            best_nt = t;
        }
    }
    nt = best_nt;
    niter = best_it;
    nvp = nt * niter;
}

// k_cells variant: registers for up to 8 velocities per thread, else the output-row round trip (0)
int nit_bucket(int niter)
{
    int b = 1;
    while (b < niter) b <<= 1;
    return b <= 8 ? b : 0;
}

double mag3(const double *a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// inverse of a symmetric tensor (xx, xy, xz, yy, yz, zz) by cofactors
void symm_inverse(const double *s, double *r)
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

int build(dugks_engine *h, const dugks_case *cs)
{
    const int nc = cs->n_cells, nfi = cs->n_internal_faces, nf = cs->n_faces, nb = nf - nfi;
    h->D = cs->dim; h->nc = nc; h->nfi = nfi; h->nf = nf; h->nb = nb;
    h->nxi = cs->n_xi; h->K = cs->K; h->R = cs->R; h->omega = cs->omega; h->Tref = cs->T_ref;
    h->mu_ref = cs->mu_ref; h->Pr = cs->Pr; h->xi_max = cs->xi_max; h->storage = cs->storage;
    h->n_cases = cs->n_cases > 1 ? cs->n_cases : 1;
    h->has_h = (cs->K + 3 - cs->dim) != 0;
    h->nv = cs->dim == 3 ? cs->n_xi * cs->n_xi * cs->n_xi : cs->n_xi * cs->n_xi;
    h->nax = cs->dim * cs->n_xi;
    const int D = h->D;

    // ---- boundary bookkeeping ----
    std::vector<int> bkind(nb, 0), bpartner(nb, -1), bpatch(nb, -1), bwall(nb, -1);
    std::vector<double> Ub(3 * (size_t)nb, 0.0), Tb(nb, 0.0);
    h->patch_type.assign(cs->n_patches, 0);
    h->patch_start.assign(cs->n_patches, 0);
    h->patch_size.assign(cs->n_patches, 0);
    h->patch_wall0.assign(cs->n_patches, -1);
    int nw = 0;
    for (int pi = 0; pi < cs->n_patches; ++pi) {
        const dugks_patch &pp = cs->patches[pi];
        if (pp.start < nfi || pp.size < 0 || pp.start + pp.size > nf)
            return fail(h, DUGKS_ERR_INVALID, "patch " + std::to_string(pi) + " face range outside the boundary");
        if (pp.type != DUGKS_PATCH_WALL && pp.type != DUGKS_PATCH_CYCLIC)
            return fail(h, DUGKS_ERR_UNSUPPORTED, "patch " + std::to_string(pi) + ": unsupported type");
        h->patch_type[pi] = pp.type; h->patch_start[pi] = pp.start; h->patch_size[pi] = pp.size;
        if (pp.type == DUGKS_PATCH_WALL) {
            if (!pp.U || !pp.T) return fail(h, DUGKS_ERR_INVALID, "wall patch without U/T");
            h->patch_wall0[pi] = nw;
        } else {
            if (pp.neighbour_patch < 0 || pp.neighbour_patch >= cs->n_patches ||
                cs->patches[pp.neighbour_patch].size != pp.size ||
                cs->patches[pp.neighbour_patch].type != DUGKS_PATCH_CYCLIC)
                return fail(h, DUGKS_ERR_INVALID, "cyclic patch without a matching partner");
        }
        for (int i = 0; i < pp.size; ++i) {
            int b = pp.start - nfi + i;
            if (bkind[b]) return fail(h, DUGKS_ERR_INVALID, "boundary face in two patches");
            bkind[b] = pp.type;
            bpatch[b] = pi;
            if (pp.type == DUGKS_PATCH_WALL) {
                for (int d = 0; d < 3; ++d) Ub[3 * b + d] = pp.U[3 * i + d];
                Tb[b] = pp.T[i];
                if (!(Tb[b] > 0)) return fail(h, DUGKS_ERR_INVALID, "wall temperature must be > 0");
                bwall[b] = nw++;
            } else {
                bpartner[b] = cs->patches[pp.neighbour_patch].start - nfi + i;
            }
        }
    }
    for (int b = 0; b < nb; ++b)
        if (!bkind[b]) return fail(h, DUGKS_ERR_INVALID, "boundary face not in any patch");
    h->nw = nw;

    // ---- mesh checks ----
    for (int f = 0; f < nf; ++f)
        if (cs->owner[f] < 0 || cs->owner[f] >= nc) return fail(h, DUGKS_ERR_INVALID, "owner out of range");
    for (int f = 0; f < nfi; ++f)
        if (cs->neighbour[f] < 0 || cs->neighbour[f] >= nc || cs->neighbour[f] == cs->owner[f])
            return fail(h, DUGKS_ERR_INVALID, "neighbour out of range");
    for (int c = 0; c < nc; ++c)
        if (!(cs->V[c] > 0) || !std::isfinite(cs->V[c]))
            return fail(h, DUGKS_ERR_INVALID, "cell " + std::to_string(c) + " has non-positive volume");
    for (int f = 0; f < nf; ++f)
        if (!(mag3(cs->Sf + 3 * f) > 0)) return fail(h, DUGKS_ERR_INVALID, "zero-area face");

    // ---- least-squares vectors (leastSquaresVectors) ----
    std::vector<double> dd(6 * (size_t)nc, 0.0), pvec(3 * (size_t)nf), nvec(3 * (size_t)nfi);
    std::vector<double> magSf(nf);
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
        if (!(d2 > 0)) return fail(h, DUGKS_ERR_INVALID, "zero boundary delta");
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
                if (!std::isfinite(r[i])) return fail(h, DUGKS_ERR_INVALID, "singular least-squares tensor");
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

    // ---- cell slots (faces ascending) ----
    std::vector<int> cnt(nc + 1, 0);
    for (int f = 0; f < nf; ++f) cnt[cs->owner[f] + 1]++;
    for (int f = 0; f < nfi; ++f) cnt[cs->neighbour[f] + 1]++;
    for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    const int nslot = cnt[nc];
    std::vector<Slot> sl(nslot);
    std::vector<int> fill(nc, 0), cface(nslot);
    for (int f = 0; f < nf; ++f) {
        int o = cs->owner[f];
        cface[cnt[o] + fill[o]++] = f;
        if (f < nfi) {
            int n = cs->neighbour[f];
            cface[cnt[n] + fill[n]++] = f;
        }
    }
    h->wall_face_slot.assign(nw, -1);
    h->ncell_slots.assign(nc, 0);
    for (int c = 0; c < nc; ++c) {
        int ns = cnt[c + 1] - cnt[c];
        h->ncell_slots[c] = ns;
        if (ns > MAXS)
            return fail(h, DUGKS_ERR_UNSUPPORTED, "cell " + std::to_string(c) + " has more than 8 faces");
        if (ns == 0) return fail(h, DUGKS_ERR_INVALID, "cell without faces");
        for (int e = cnt[c]; e < cnt[c + 1]; ++e) {
            const int f = cface[e];
            Slot s;
            std::memset(&s, 0, sizeof s);
            s.face = f;
            for (int d = 0; d < 3; ++d) {
                s.S[d] = cs->Sf[3 * f + d];
                s.dCs[d] = cs->Cf[3 * f + d] - Cc[3 * c + d];
            }
            s.bidx = -1;
            s.other = -1;
            if (f < nfi) {
                const bool own = cs->owner[f] == c;
                s.kind = own ? KIND_OWN : KIND_NEI;
                s.other = own ? cs->neighbour[f] : cs->owner[f];
                for (int d = 0; d < 3; ++d) {
                    s.L[d] = own ? pvec[3 * f + d] : nvec[3 * f + d];
                    s.dCo[d] = cs->Cf[3 * f + d] - Cc[3 * s.other + d];
                }
            } else {
                const int b = f - nfi;
                for (int d = 0; d < 3; ++d) s.L[d] = pvec[3 * f + d];
                if (bkind[b] == DUGKS_PATCH_WALL) {
                    s.kind = KIND_WALL;
                    s.bidx = bwall[b];
                    for (int d = 0; d < 3; ++d) s.dCo[d] = cs->Sf[3 * f + d] / magSf[f];
                    const double *dl = cs->delta + 3 * b;
                    s.aux = 1.0 / (1.0 / std::sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]));
                    h->wall_face_slot[bwall[b]] = e;
                } else {
                    s.kind = KIND_CYC;
                    s.bidx = b;
                    s.other = cs->owner[nfi + bpartner[b]];
                    s.aux = w[f];
                }
            }
            sl[e] = s;
        }
    }
    // ---- per-face data of k_fmom ----
    std::vector<FaceInfo> finfo(nf);
    for (int f = 0; f < nf; ++f) {
        FaceInfo fi;
        std::memset(&fi, 0, sizeof fi);
        for (int d = 0; d < 3; ++d) fi.S[d] = cs->Sf[3 * f + d];
        fi.own = cs->owner[f];
        fi.kind = KIND_OWN;
        fi.bidx = -1;
        if (f >= nfi) {
            const int b = f - nfi;
            fi.kind = bkind[b] == DUGKS_PATCH_WALL ? KIND_WALL : KIND_CYC;
            fi.bidx = bkind[b] == DUGKS_PATCH_WALL ? bwall[b] : b;
        }
        finfo[f] = fi;
    }

    // ---- velocities: tensor product in reference order, then grouped by sign pattern ----
    const int nxi = h->nxi;
    if (nxi < 2 || nxi > 64) return fail(h, DUGKS_ERR_UNSUPPORTED, "n_xi must be in [2, 64]");
    if (h->nax > NAXMAX) return fail(h, DUGKS_ERR_UNSUPPORTED, "velocity table too large");
    const int nv = h->nv;
    std::vector<double> rx(nv), ry(nv), rz(nv), rw(nv);
    std::vector<int> rid(nv);
    {
        int i = 0;
        const double *X = cs->xi, *W = cs->xi_weights;
        if (D == 3) {
            for (int iz = 0; iz < nxi; ++iz)
                for (int iy = 0; iy < nxi; ++iy)
                    for (int ix = 0; ix < nxi; ++ix) {
                        rw[i] = W[iz] * W[iy] * W[ix];
                        rx[i] = X[ix]; ry[i] = X[iy]; rz[i] = X[iz];
                        rid[i] = ix | (iy << 8) | (iz << 16);
                        ++i;
                    }
        } else {
            for (int iy = 0; iy < nxi; ++iy)
                for (int ix = 0; ix < nxi; ++ix) {
                    rw[i] = W[iy] * W[ix] * 1;
                    rx[i] = X[ix]; ry[i] = X[iy]; rz[i] = 0.0;
                    rid[i] = ix | (iy << 8);
                    ++i;
                }
        }
    }
    auto cls = [](double x) { return x < 0 ? 0 : (x == 0 ? 1 : 2); };
    std::vector<int> order(nv);
    for (int i = 0; i < nv; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        int ka = cls(rz[a]) * 9 + cls(ry[a]) * 3 + cls(rx[a]);
        int kb = cls(rz[b]) * 9 + cls(ry[b]) * 3 + cls(rx[b]);
        return ka < kb;
    });
    choose_launch(nv, h->nt, h->niter, h->nvp);
    if (h->nt <= 0 || h->nt > MAXNT || h->niter > 64 || h->nt * h->niter < nv)
        return fail(h, DUGKS_ERR_UNSUPPORTED, "velocity count too large for one workgroup per cell");
    h->nit_max = nit_bucket(h->niter);
    {
        int b = 1;
        while (b < h->niter) b <<= 1;
        h->fmom_nit = b <= 32 ? b : 0;
    }
    const int nvp = h->nvp;
    std::vector<VelRec> vel(nvp);
    std::memset(vel.data(), 0, sizeof(VelRec) * nvp);
    h->perm.assign(nvp, -1);
    for (int k = 0; k < nv; ++k) {
        const int r = order[k];
        vel[k].x = rx[r]; vel[k].y = ry[r]; vel[k].z = rz[r]; vel[k].w = rw[r];
        vel[k].xi2 = rx[r] * rx[r] + ry[r] * ry[r] + rz[r] * rz[r];
        vel[k].idx = rid[r];
        h->perm[k] = r;
    }
    std::vector<double> axis(h->nax);
    for (int a = 0; a < D; ++a)
        for (int i = 0; i < nxi; ++i) axis[a * nxi + i] = cs->xi[i];

    // ---- per-case data ----
    std::vector<int> ccase(nc, 0);
    std::vector<double> muc(h->n_cases, cs->mu_ref);
    if (h->n_cases > 1) {
        if (!cs->cell_case) return fail(h, DUGKS_ERR_INVALID, "n_cases > 1 without cell_case");
        for (int c = 0; c < nc; ++c) {
            if (cs->cell_case[c] < 0 || cs->cell_case[c] >= h->n_cases)
                return fail(h, DUGKS_ERR_INVALID, "cell_case out of range");
            ccase[c] = cs->cell_case[c];
        }
        if (cs->mu_ref_case)
            for (int i = 0; i < h->n_cases; ++i) muc[i] = cs->mu_ref_case[i];
        for (int f = 0; f < nfi; ++f)
            if (ccase[cs->owner[f]] != ccase[cs->neighbour[f]])
                return fail(h, DUGKS_ERR_INVALID, "a face joins two cases");
    }
    h->face_w.assign(w, w + nf);
    h->dcoef_int.resize(nfi);
    for (int f = 0; f < nfi; ++f) {
        int o = cs->owner[f], n = cs->neighbour[f];
        double d[3] = {Cc[3 * n] - Cc[3 * o], Cc[3 * n + 1] - Cc[3 * o + 1], Cc[3 * n + 2] - Cc[3 * o + 2]};
        h->dcoef_int[f] = 1.0 / mag3(d);
    }

    // ---- wall constants (U, T, pre); the sums come from k_wall_init ----
    std::vector<WallC> wc(nw);
    std::vector<int> wfslot(nw);
    for (int b = 0; b < nb; ++b) {
        if (bwall[b] < 0) continue;
        WallC W;
        std::memset(&W, 0, sizeof W);
        for (int d = 0; d < 3; ++d) W.U[d] = Ub[3 * b + d];
        W.T = Tb[b];
        double tp = 2.0 * PI * (cs->R * Tb[b]);
        W.pre = D == 3 ? 1.0 / (tp * std::sqrt(tp)) : 1.0 / tp;
        wc[bwall[b]] = W;
    }
    for (int i = 0; i < nw; ++i) {
        wfslot[i] = h->wall_face_slot[i];
        if (wfslot[i] < 0) return fail(h, DUGKS_ERR_INVALID, "wall face without slot");
    }

    // ---- device allocation and upload ----
    HIPCHK(h, hipSetDevice(cs->device));
    h->device = cs->device;
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    int r = 0;
    if ((r = dalloc(h, &h->vel, nvp)) ||
        (r = dalloc(h, &h->axis, h->nax)) || (r = dalloc(h, &h->cs, nc + 1)) || (r = dalloc(h, &h->slots, nslot)) ||
        (r = dalloc(h, &h->cV, nc)) || (r = dalloc(h, &h->ccase, nc)) || (r = dalloc(h, &h->dtc, h->n_cases)) ||
        (r = dalloc(h, &h->dtp, h->n_cases)) || (r = dalloc(h, &h->muc, h->n_cases)) ||
        (r = dalloc(h, &h->crho, nc)) || (r = dalloc(h, &h->cU, 3 * (size_t)nc)) || (r = dalloc(h, &h->cT, nc)) ||
        (r = dalloc(h, &h->cq, 3 * (size_t)nc)) || (r = dalloc(h, &h->ctau, nc)) || (r = dalloc(h, &h->frho, nf)) ||
        (r = dalloc(h, &h->fU, 3 * (size_t)nf)) || (r = dalloc(h, &h->fT, nf)) || (r = dalloc(h, &h->fq, 3 * (size_t)nf)) ||
        (r = dalloc(h, &h->feq, nf)) ||
        (r = dalloc(h, &h->wall, nw)) || (r = dalloc(h, &h->wexp, (size_t)nw * h->nax)) ||
        (r = dalloc(h, &h->qwall, 3 * (size_t)nw)) || (r = dalloc(h, &h->swall, 9 * (size_t)nw)) ||
        (r = dalloc(h, &h->finfo, nf)) || (r = dalloc(h, &h->flag, 1)) || (r = dalloc(h, &h->wfslot, nw)))
        return r;
    const size_t es = ts_size(h);
    const size_t row = (size_t)nc * nvp * es, brow = (size_t)(nw ? nw : 1) * nvp * es;
    const size_t frow = (size_t)nf * nvp * es;
    if ((r = dallocb(h, &h->fg, frow))) return r;
    if (h->has_h && (r = dallocb(h, &h->fh, frow))) return r;
    for (int i = 0; i < 2; ++i) {
        if ((r = dallocb(h, &h->gb[i], row))) return r;
        if (h->has_h && (r = dallocb(h, &h->hb[i], row))) return r;
        if ((r = dallocb(h, &h->gB[i], brow))) return r;
        if (h->has_h && (r = dallocb(h, &h->hB[i], brow))) return r;
    }
    if ((r = upload(h, h->vel, vel.data(), nvp)) ||
        (r = upload(h, h->axis, axis.data(), h->nax)) || (r = upload(h, h->cs, cnt.data(), nc + 1)) ||
        (r = upload(h, h->slots, sl.data(), nslot)) || (r = upload(h, h->cV, cs->V, nc)) ||
        (r = upload(h, h->ccase, ccase.data(), nc)) || (r = upload(h, h->muc, muc.data(), h->n_cases)) ||
        (r = upload(h, h->finfo, finfo.data(), nf)) || (r = upload(h, h->wfslot, wfslot.data(), nw)))
        return r;
    if (nw && (r = upload(h, h->wall, wc.data(), nw))) return r;
    HIPCHK(h, hipMemset(h->flag, 0, sizeof(int)));
    h->dt_prep.assign(h->n_cases, 0.0);
    return 0;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

void release(dugks_engine *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    dfree(h->vel); dfree(h->axis);
    dfree(h->cs); dfree(h->ccase); dfree(h->flag); dfree(h->wfslot); dfree(h->slots); dfree(h->cV);
    dfree(h->dtc); dfree(h->dtp); dfree(h->muc); dfree(h->crho); dfree(h->cU); dfree(h->cT); dfree(h->cq);
    dfree(h->ctau); dfree(h->frho); dfree(h->fU); dfree(h->fT); dfree(h->fq); dfree(h->feq);
    dfree(h->wall); dfree(h->wexp); dfree(h->qwall); dfree(h->swall); dfree(h->finfo);
    dfree(h->fg); dfree(h->fh);
    for (int i = 0; i < 2; ++i) {
        dfree(h->gb[i]); dfree(h->hb[i]); dfree(h->gB[i]); dfree(h->hB[i]);
    }
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    for (hipEvent_t e : h->pev) (void)hipEventDestroy(e);
    h->pev.clear();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->ev0 = h->ev1 = nullptr;
    h->stream = nullptr;
    h->initialised = 0;
}

int check_flag(dugks_engine *h)
{
    int f = 0;
    HIPCHK(h, hipMemcpyAsync(&f, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (f) {
        HIPCHK(h, hipMemset(h->flag, 0, sizeof(int)));
        return fail(h, DUGKS_ERR_NONFINITE, "step produced non-finite or non-positive density/temperature");
    }
    return 0;
}

int step_impl(dugks_engine *h, const double *dtc, int nsteps)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (!h->initialised) return fail(h, DUGKS_ERR_STATE, "dugks_step before dugks_init_equilibrium");
    if (nsteps < 0) return fail(h, DUGKS_ERR_INVALID, "nsteps < 0");
    for (int i = 0; i < h->n_cases; ++i)
        if (!(dtc[i] > 0) || !std::isfinite(dtc[i])) return fail(h, DUGKS_ERR_INVALID, "dt must be > 0");
    HIPCHK(h, hipSetDevice(h->device));
    h->dt_host.assign(dtc, dtc + h->n_cases);
    HIPCHK(h, hipMemcpyAsync(h->dtc, h->dt_host.data(), sizeof(double) * h->n_cases, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    // bring the resident state to gbar prepared with this dt
    bool same = !h->raw;
    for (int i = 0; i < h->n_cases && same; ++i) same = h->dt_prep[i] == dtc[i];
    if (!same && nsteps > 0) {
        ConvertOp op{h, h->raw ? 0 : 1};
        dispatch(h, op);
        h->cur = 1 - h->cur;
        HIPCHK(h, hipGetLastError());
        h->raw = 0;
        for (int i = 0; i < h->n_cases; ++i) h->dt_prep[i] = dtc[i];
        HIPCHK(h, hipMemcpyAsync(h->dtp, h->dt_host.data(), sizeof(double) * h->n_cases, hipMemcpyHostToDevice, h->stream));
    }
    if (h->profiling) {
        while ((int)h->pev.size() < 4 * nsteps) {
            hipEvent_t e;
            HIPCHK(h, hipEventCreate(&e));
            h->pev.push_back(e);
        }
    }
    for (int s = 0; s < nsteps; ++s) {
        StepOp op{h, h->profiling ? &h->pev[4 * (size_t)s] : nullptr};
        dispatch(h, op);
        h->cur = 1 - h->cur;
        h->gcur = 1 - h->gcur;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    int r = check_flag(h);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->last_ms = ms;
    if (h->profiling) {
        for (int s = 0; s < nsteps; ++s) {
            float a = 0.f, b = 0.f, c = 0.f;
            const hipEvent_t *e = &h->pev[4 * (size_t)s];
            HIPCHK(h, hipEventElapsedTime(&a, e[0], e[1]));
            HIPCHK(h, hipEventElapsedTime(&b, e[1], e[2]));
            HIPCHK(h, hipEventElapsedTime(&c, e[2], e[3]));
            h->prof_recon += a;
            h->prof_fmom += b;
            h->prof_cells += c;
        }
        h->prof_steps += nsteps;
    }
    return r;
}

const int64_t STATE_MAGIC = 0x44554b5331LL;  // "DUKS1"

}  // namespace

extern "C" {

int dugks_abi_version(void) { return DUGKS_ABI_VERSION; }

int dugks_create(const dugks_case *cs, dugks_engine **out)
{
    if (!out) return DUGKS_ERR_INVALID;
    *out = nullptr;
    if (!cs) return DUGKS_ERR_INVALID;
    dugks_engine *h = new (std::nothrow) dugks_engine();
    if (!h) return DUGKS_ERR_MEMORY;
    int r = 0;
    if (cs->abi_version != DUGKS_ABI_VERSION) r = fail(h, DUGKS_ERR_INVALID, "abi_version mismatch");
    else if (cs->dim != 2 && cs->dim != 3) r = fail(h, DUGKS_ERR_UNSUPPORTED, "dim must be 2 or 3");
    else if (cs->n_cells <= 0 || cs->n_internal_faces < 0 || cs->n_faces < cs->n_internal_faces)
        r = fail(h, DUGKS_ERR_INVALID, "mesh sizes");
    else if (!cs->owner || (cs->n_internal_faces && !cs->neighbour) || !cs->C || !cs->V || !cs->Sf || !cs->Cf ||
             !cs->weights || (cs->n_faces > cs->n_internal_faces && !cs->delta) || !cs->xi || !cs->xi_weights ||
             (cs->n_patches && !cs->patches))
        r = fail(h, DUGKS_ERR_INVALID, "null mesh array");
    else if (!(cs->R > 0) || !(cs->T_ref > 0) || !(cs->mu_ref > 0) || !(cs->Pr > 0) || cs->K < 0)
        r = fail(h, DUGKS_ERR_INVALID, "gas properties");
    else if (cs->storage != DUGKS_STORAGE_FP64 && cs->storage != DUGKS_STORAGE_FP32)
        r = fail(h, DUGKS_ERR_INVALID, "storage");
    else {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) r = fail(h, DUGKS_ERR_DEVICE, "no HIP device");
        else if (cs->device < 0 || cs->device >= ndev) r = fail(h, DUGKS_ERR_INVALID, "device ordinal");
        else r = build(h, cs);
    }
    if (r) {
        // the handle is returned without device resources so that dugks_last_error can
        // report the reason; the caller still calls dugks_destroy
        release(h);
        *out = h;
        return r;
    }
    *out = h;
    return 0;
}

int dugks_init_equilibrium(dugks_engine *h, const double *rho, const double *U, const double *T)
{
    if (!h || !h->stream) return DUGKS_ERR_STATE;
    if (!rho || !U || !T) return fail(h, DUGKS_ERR_INVALID, "null field");
    const int nc = h->nc, nf = h->nf, nfi = h->nfi;
    for (int c = 0; c < nc; ++c)
        if (!(rho[c] > 0) || !(T[c] > 0) || !std::isfinite(rho[c]) || !std::isfinite(T[c]) ||
            !std::isfinite(U[3 * c]) || !std::isfinite(U[3 * c + 1]) || !std::isfinite(U[3 * c + 2]))
            return fail(h, DUGKS_ERR_INVALID, "initial field invalid at cell " + std::to_string(c));
    HIPCHK(h, hipSetDevice(h->device));
    std::vector<double> tau(nc), q(3 * (size_t)nc, 0.0);
    std::vector<int> ccase(nc, 0);
    HIPCHK(h, hipMemcpy(ccase.data(), h->ccase, sizeof(int) * nc, hipMemcpyDeviceToHost));
    std::vector<double> muc(h->n_cases);
    HIPCHK(h, hipMemcpy(muc.data(), h->muc, sizeof(double) * h->n_cases, hipMemcpyDeviceToHost));
    for (int c = 0; c < nc; ++c) tau[c] = muc[ccase[c]] * std::exp(h->omega * std::log(T[c] / h->Tref)) / rho[c] / T[c] / h->R;
    int r;
    if ((r = upload(h, h->crho, rho, nc)) || (r = upload(h, h->cU, U, 3 * (size_t)nc)) ||
        (r = upload(h, h->cT, T, nc)) || (r = upload(h, h->cq, q.data(), 3 * (size_t)nc)) ||
        (r = upload(h, h->ctau, tau.data(), nc)))
        return r;
    // Usurf = linear interpolate(U) (fvDVM.C:1074), the reference velocity of the first face moments
    std::vector<double> fU(3 * (size_t)nf, 0.0);
    std::vector<Slot> sl;
    {
        std::vector<int> cs(nc + 1);
        HIPCHK(h, hipMemcpy(cs.data(), h->cs, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost));
        sl.resize(cs[nc]);
        HIPCHK(h, hipMemcpy(sl.data(), h->slots, sizeof(Slot) * cs[nc], hipMemcpyDeviceToHost));
        std::vector<WallC> wc(h->nw);
        if (h->nw) HIPCHK(h, hipMemcpy(wc.data(), h->wall, sizeof(WallC) * h->nw, hipMemcpyDeviceToHost));
        for (int c = 0; c < nc; ++c)
            for (int e = cs[c]; e < cs[c + 1]; ++e) {
                const Slot &s = sl[e];
                if (s.kind == KIND_NEI) continue;
                const int f = s.face;
                if (s.kind == KIND_WALL) {
                    for (int d = 0; d < 3; ++d) fU[3 * f + d] = wc[s.bidx].U[d];
                } else {
                    // linear interpolation w U_P + (1 - w) U_N (cyclic: patch weight)
                    const double wf = h->face_w[f];
                    for (int d = 0; d < 3; ++d) fU[3 * f + d] = wf * U[3 * c + d] + (1 - wf) * U[3 * s.other + d];
                }
            }
    }
    (void)nfi;
    if ((r = upload(h, h->fU, fU.data(), 3 * (size_t)nf))) return r;
    h->cur = 0;
    h->gcur = 0;
    const size_t es = ts_size(h);
    const size_t brow = (size_t)(h->nw ? h->nw : 1) * h->nvp * es;
    for (int i = 0; i < 2; ++i) {
        HIPCHK(h, hipMemset(h->gB[i], 0, brow));
        if (h->has_h) HIPCHK(h, hipMemset(h->hB[i], 0, brow));
    }
    InitOp op{h};
    dispatch(h, op);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->raw = 1;
    h->initialised = 1;
    return 0;
}

int dugks_step(dugks_engine *h, double dt, int32_t nsteps)
{
    if (!h || !h->stream) return DUGKS_ERR_STATE;
    std::vector<double> d(h->n_cases, dt);
    return step_impl(h, d.data(), nsteps);
}

int dugks_step_cases(dugks_engine *h, const double *dt_case, int32_t nsteps)
{
    if (!h || !h->stream) return DUGKS_ERR_STATE;
    if (!dt_case) return fail(h, DUGKS_ERR_INVALID, "null dt_case");
    return step_impl(h, dt_case, nsteps);
}

int dugks_get_macro(dugks_engine *h, double *rho, double *U, double *T, double *q)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const size_t nc = h->nc;
    if (rho) HIPCHK(h, hipMemcpy(rho, h->crho, sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (U) HIPCHK(h, hipMemcpy(U, h->cU, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost));
    if (T) HIPCHK(h, hipMemcpy(T, h->cT, sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (q) HIPCHK(h, hipMemcpy(q, h->cq, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost));
    return 0;
}

int dugks_get_face_macro(dugks_engine *h, double *rho, double *U, double *T, double *q)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const size_t nf = h->nf;
    if (rho) HIPCHK(h, hipMemcpy(rho, h->frho, sizeof(double) * nf, hipMemcpyDeviceToHost));
    if (U) HIPCHK(h, hipMemcpy(U, h->fU, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost));
    if (T) HIPCHK(h, hipMemcpy(T, h->fT, sizeof(double) * nf, hipMemcpyDeviceToHost));
    if (q) HIPCHK(h, hipMemcpy(q, h->fq, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost));
    return 0;
}

int dugks_get_wall(dugks_engine *h, int32_t patch, double *rho_w, double *q_wall, double *stress_wall)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    if (patch < 0 || patch >= (int)h->patch_type.size() || h->patch_type[patch] != DUGKS_PATCH_WALL)
        return fail(h, DUGKS_ERR_INVALID, "not a wall patch");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int w0 = h->patch_wall0[patch], n = h->patch_size[patch], f0 = h->patch_start[patch];
    if (rho_w) {
        std::vector<FaceEq> fe(n);
        HIPCHK(h, hipMemcpy(fe.data(), h->feq + f0, sizeof(FaceEq) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) rho_w[i] = fe[i].rho_w;
    }
    if (q_wall) HIPCHK(h, hipMemcpy(q_wall, h->qwall + 3 * (size_t)w0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    if (stress_wall)
        HIPCHK(h, hipMemcpy(stress_wall, h->swall + 9 * (size_t)w0, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
    return 0;
}

int dugks_get_courant(dugks_engine *h, double dt, double *co_max, double *co_mean)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    if (!co_max || !co_mean) return fail(h, DUGKS_ERR_INVALID, "null output");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> fU(3 * (size_t)h->nfi);
    if (h->nfi) HIPCHK(h, hipMemcpy(fU.data(), h->fU, sizeof(double) * 3 * h->nfi, hipMemcpyDeviceToHost));
    double mx = -1e300, sum = 0.0;
    const double sd = std::sqrt((double)h->D) * h->xi_max;
    for (int f = 0; f < h->nfi; ++f) {
        double u = h->dcoef_int[f] * (mag3(&fU[3 * (size_t)f]) + sd);
        mx = std::max(mx, u);
        sum += u;
    }
    *co_max = mx * dt;
    *co_mean = h->nfi ? sum / h->nfi * dt : 0.0;
    return 0;
}

int dugks_get_df(dugks_engine *h, int32_t n, const int32_t *cells, double *g, double *h_out)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    if (n < 0 || (n && !cells)) return fail(h, DUGKS_ERR_INVALID, "cells");
    for (int i = 0; i < n; ++i)
        if (cells[i] < 0 || cells[i] >= h->nc) return fail(h, DUGKS_ERR_INVALID, "cell index out of range");
    if (n == 0) return 0;
    HIPCHK(h, hipSetDevice(h->device));
    int *dcells = nullptr;
    double *dg = nullptr, *dh = nullptr;
    int r = 0;
    if ((r = dalloc(h, &dcells, n)) || (r = dalloc(h, &dg, (size_t)n * h->nvp)) || (r = dalloc(h, &dh, (size_t)n * h->nvp))) {
        if (dcells) (void)hipFree(dcells);
        if (dg) (void)hipFree(dg);
        return r;
    }
    std::vector<double> hg((size_t)n * h->nvp), hh((size_t)n * h->nvp);
    hipError_t e = hipMemcpyAsync(dcells, cells, sizeof(int) * n, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        ExtractOp op{h, dcells, n, dg, dh};
        dispatch(h, op);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hg.data(), dg, sizeof(double) * hg.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hh.data(), dh, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dcells);
    (void)hipFree(dg);
    (void)hipFree(dh);
    if (e != hipSuccess) return fail(h, DUGKS_ERR_DEVICE, std::string("get_df: ") + hipGetErrorString(e));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < h->nvp; ++k) {
            const int rk = h->perm[k];
            if (rk < 0) continue;
            if (g) g[(size_t)i * h->nv + rk] = hg[(size_t)i * h->nvp + k];
            if (h_out) h_out[(size_t)i * h->nv + rk] = hh[(size_t)i * h->nvp + k];
        }
    return 0;
}

int dugks_get_sizes(dugks_engine *h, int32_t *nv, int32_t *n_cells, int32_t *n_faces)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (nv) *nv = h->nv;
    if (n_cells) *n_cells = h->nc;
    if (n_faces) *n_faces = h->nf;
    return 0;
}

// State image: header, gbar/hbar (current buffer), boundary gradients (current), cell macros,
// face reference velocities, prepared dt per case.
int dugks_state_size(dugks_engine *h, int64_t *bytes)
{
    if (!h || !bytes) return DUGKS_ERR_INVALID;
    const int64_t es = (int64_t)ts_size(h);
    const int64_t nh = h->has_h ? 2 : 1;
    int64_t n = 16 * 8;
    n += nh * (int64_t)h->nc * h->nvp * es;
    n += nh * (int64_t)(h->nw ? h->nw : 1) * h->nvp * es;
    n += 9 * (int64_t)h->nc * 8 + 3 * (int64_t)h->nf * 8 + (int64_t)h->n_cases * 8;
    *bytes = n;
    return 0;
}

int dugks_save_state(dugks_engine *h, void *buf, int64_t bytes)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    int64_t need = 0;
    dugks_state_size(h, &need);
    if (!buf || bytes < need) return fail(h, DUGKS_ERR_INVALID, "state buffer too small");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    char *p = (char *)buf;
    int64_t hdr[16] = {STATE_MAGIC, h->nc, h->nvp, h->nw, h->nf, h->storage, h->has_h, h->raw, h->n_cases, h->D, 0, 0, 0, 0, 0, 0};
    std::memcpy(p, hdr, sizeof hdr);
    p += sizeof hdr;
    const size_t es = ts_size(h);
    const size_t row = (size_t)h->nc * h->nvp * es, brow = (size_t)(h->nw ? h->nw : 1) * h->nvp * es;
    HIPCHK(h, hipMemcpy(p, h->gb[h->cur], row, hipMemcpyDeviceToHost)); p += row;
    if (h->has_h) { HIPCHK(h, hipMemcpy(p, h->hb[h->cur], row, hipMemcpyDeviceToHost)); p += row; }
    HIPCHK(h, hipMemcpy(p, h->gB[h->gcur], brow, hipMemcpyDeviceToHost)); p += brow;
    if (h->has_h) { HIPCHK(h, hipMemcpy(p, h->hB[h->gcur], brow, hipMemcpyDeviceToHost)); p += brow; }
    const size_t nc = h->nc;
    HIPCHK(h, hipMemcpy(p, h->crho, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->cU, 24 * nc, hipMemcpyDeviceToHost)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(p, h->cT, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->cq, 24 * nc, hipMemcpyDeviceToHost)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(p, h->ctau, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->fU, 24 * (size_t)h->nf, hipMemcpyDeviceToHost)); p += 24 * (size_t)h->nf;
    std::memcpy(p, h->dt_prep.data(), 8 * h->n_cases);
    return 0;
}

int dugks_load_state(dugks_engine *h, const void *buf, int64_t bytes)
{
    if (!h || !h->stream) return DUGKS_ERR_STATE;
    if (!h->initialised) return fail(h, DUGKS_ERR_STATE, "call dugks_init_equilibrium before loading a state");
    int64_t need = 0;
    dugks_state_size(h, &need);
    if (!buf || bytes < need) return fail(h, DUGKS_ERR_INVALID, "state buffer too small");
    const char *p = (const char *)buf;
    int64_t hdr[16];
    std::memcpy(hdr, p, sizeof hdr);
    if (hdr[0] != STATE_MAGIC || hdr[1] != h->nc || hdr[2] != h->nvp || hdr[3] != h->nw || hdr[4] != h->nf ||
        hdr[5] != h->storage || hdr[6] != (int64_t)h->has_h || hdr[8] != h->n_cases || hdr[9] != h->D)
        return fail(h, DUGKS_ERR_INVALID, "state image does not match this engine");
    p += sizeof hdr;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const size_t es = ts_size(h);
    const size_t row = (size_t)h->nc * h->nvp * es, brow = (size_t)(h->nw ? h->nw : 1) * h->nvp * es;
    HIPCHK(h, hipMemcpy(h->gb[h->cur], p, row, hipMemcpyHostToDevice)); p += row;
    if (h->has_h) { HIPCHK(h, hipMemcpy(h->hb[h->cur], p, row, hipMemcpyHostToDevice)); p += row; }
    HIPCHK(h, hipMemcpy(h->gB[h->gcur], p, brow, hipMemcpyHostToDevice)); p += brow;
    if (h->has_h) { HIPCHK(h, hipMemcpy(h->hB[h->gcur], p, brow, hipMemcpyHostToDevice)); p += brow; }
    const size_t nc = h->nc;
    HIPCHK(h, hipMemcpy(h->crho, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->cU, p, 24 * nc, hipMemcpyHostToDevice)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(h->cT, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->cq, p, 24 * nc, hipMemcpyHostToDevice)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(h->ctau, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->fU, p, 24 * (size_t)h->nf, hipMemcpyHostToDevice)); p += 24 * (size_t)h->nf;
    std::memcpy(h->dt_prep.data(), p, 8 * h->n_cases);
    h->raw = (int)hdr[7];
    HIPCHK(h, hipMemcpy(h->dtp, h->dt_prep.data(), 8 * h->n_cases, hipMemcpyHostToDevice));
    return 0;
}

int dugks_last_step_ms(dugks_engine *h, double *ms)
{
    if (!h || !ms) return DUGKS_ERR_INVALID;
    *ms = h->last_ms;
    return 0;
}

int dugks_set_profiling(dugks_engine *h, int32_t on)
{
    if (!h) return DUGKS_ERR_INVALID;
    h->profiling = on ? 1 : 0;
    h->prof_recon = h->prof_fmom = h->prof_cells = 0.0;
    h->prof_steps = 0;
    return 0;
}

int dugks_get_profile(dugks_engine *h, double *ms, int64_t *steps)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (ms) {
        ms[0] = h->prof_recon;
        ms[1] = h->prof_fmom;
        ms[2] = h->prof_cells;
    }
    if (steps) *steps = h->prof_steps;
    return 0;
}

void dugks_destroy(dugks_engine *h)
{
    if (!h) return;
    release(h);
    delete h;
}

const char *dugks_last_error(dugks_engine *h)
{
    if (!h) return "null engine";
    return h->err.c_str();
}

}  // extern "C"
