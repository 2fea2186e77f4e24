// dugks_engine.hip -- MI355X (gfx950) DUGKS time-step engine behind the C-ABI of include/dugks.h.
//
// One time step of dugksFoam's fvDVM::evolution() (src/fvDVM/fvDVM/fvDVM.C:1086-1108) as two
// kernels over resident device state, separated by the step's two velocity-space reductions:
//
//   k_face   (one workgroup per face, all velocities):
//            S2 upwind reconstruction of the face value from the resident gBarPvol of the upwind
//            cell and its least-squares gradient, S3/S4 Maxwell wall (outGoing, rho_w), S6 face
//            moments reduced across the velocity lanes; writes the face macros, the S7 equilibrium
//            constants and the lagged fixedGradient boundary gradient of the next step.
//   k_cell   (one workgroup per cell, all velocities):
//            rebuilds the upwind value of each of its faces the same way, S7 relaxation, S8 flux
//            divergence and g~ update in the reference's per-cell face order, S9 cell moments and
//            macros, then S1 of the next step (gBarPvol = (1-r) g~ + r geq), written to the other
//            ping-pong buffer.
//
// No per-(face, velocity) array exists: face values are recomputed from the resident state where
// they are needed, so a step streams the distributions through HBM three times (two reads, one
// write) and the neighbour rows of the gradients come from L2.
//
// Data layout: distributions are [cell][velocity] with the velocity index contiguous and padded to
// NVP = NT*NITER; a workgroup's lanes are velocities of one cell or face, so every mesh quantity is
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
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "lsq.h"

namespace {

constexpr double VSMALL = 1.0e-300;   // OpenFOAM VSMALL (doubleScalar.H)
constexpr double PI = 3.14159265358979323846;
constexpr int KIND_OWN = 0;   // internal face, this cell is the owner
constexpr int KIND_NEI = 1;   // internal face, this cell is the neighbour
constexpr int KIND_WALL = 2;  // Maxwell wall (fixedGradient gBarPvol, maxwellWall gSurf)
constexpr int KIND_CYC = 3;   // cyclic (coupled) patch face
constexpr int KIND_SYM = 4;   // symmetry plane (symmetryMod rho, DVMsymmetry gSurf, fixedGradient gBarPvol)
constexpr int KIND_FAR = 5;   // incoming Maxwellian: far field (farField rho) or fixed (fixedValue rho)
constexpr int KIND_ZG = 6;    // zeroGradient: every face value is the owner cell's gBarPvol
constexpr int KIND_SPL = 7;   // OpenFOAM symmetryPlane: mirror plane, gBarPvol phi_b = phi_P (no lagged
                              // row), every value relaxed by the field operation, outgoing ones twice

// Boundary faces with a lagged fixedGradient gBarPvol row; open faces relax only outgoing values.
__host__ __device__ constexpr inline bool grad_row(int kind)
{
    return kind == KIND_WALL || kind == KIND_SYM || kind == KIND_FAR || kind == KIND_ZG;
}
constexpr int MAXS = 16;      // faces per cell (stencil templates NS = 4, 6, 8, 16)
constexpr int NAXMAX = 192;   // per-axis exp table entries (3 axes x 64 points)
constexpr int MAXNT = 512;     // workgroup size cap (256 VGPRs per lane available)
constexpr int MAXWAVES = MAXNT / 64;
constexpr int NIBMAX = 64;     // velocity iterations per thread in a block
constexpr int S1_UNR = 4;      // velocities per thread in flight in k_cell_s1
// Minimum resident waves per SIMD asked of the register allocator (measured on the 48^3 x 24^3
// cavity: 6 for the FP64 k_face costs a few spill bytes and gains 4 %, the FP32 one spills more
// and is held at 5; 5 for k_cell spills in its face loop and loses 35 %; the face-value k_cell,
// which rebuilds no gradients, runs best at 5 of 4, 5, 6, 8).
#ifndef FACE_WAVES
#define FACE_WAVES 6
#endif
#ifndef FACE_WAVES_FP32
#define FACE_WAVES_FP32 5
#endif
// k_face_values asks for 8 waves where that costs no spills: g-only stencils of up to 8 faces and
// the 2-D and 1-D stencils of 4.  The register allocator then also keeps the kernel under 100
// SGPRs; at the lower target it takes 106 and the SGPR file caps the SIMD at 7 waves (the
// 48^3 x 24^3 cavity's face pass: 14.0-14.4 instead of 14.8-15.2 ms in FP64).
#ifndef KFV_WAVES
#define KFV_WAVES 8
#endif
#ifndef CELL_WAVES
#define CELL_WAVES 4
#endif
#ifndef CELL_WAVES_FV
#define CELL_WAVES_FV 5
#endif
#ifndef KC_PRELOAD
#define KC_PRELOAD 1
#endif
// cells whose faces are all internal take a loop without the boundary branches
#ifndef KFV_INNER
#define KFV_INNER 1
#endif
#ifndef FM_UNR
#define FM_UNR 2
#endif

// One (cell, face) incidence; a cell's slots are in ascending face order, the order in which the
// reference's face loops touch that cell.
struct alignas(16) Slot {
    double L[3];    // least-squares vector of this cell for this face
    double S[3];    // face area vector
    double dCs[3];  // Cf - C(this cell)
    double dCo[3];  // Cf - C(other cell) | wall, symmetry: unit normal Sf/|Sf|
    double aux;     // wall, symmetry: 1/deltaCoeffs | cyclic: patch weight w
    int32_t face, other, kind, bidx;  // other: the other cell | symmetry: mirror axis
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
    double pad[5];  // walls: [0] T_w, [1] 1 / (2 pi R T_w)^(D/2)
    // the equilibrium polynomial in xi: c2 - D - 2 = |xi|^2 invRT + B.xi + C0 and
    // pc = -(U - xi).qc = Q0 + qc.xi, with c2 = |U - xi|^2 / RT
    double B[3];    // -2 U / RT
    double C0;      // |U|^2 / RT - (D + 2)
    double Q0;      // -U.qc
    double omr;     // 1 - r
    double pad2[2];
};
static_assert(sizeof(FaceEq) == 192, "faceeq size");

// Wall face constants: wall velocity/temperature, inComingByRho and the moments of the incoming
// Maxwellian (per unit rho_w) about the wall velocity.
struct alignas(16) WallC {
    double U[3];
    double T;
    double incoming;
    double pre;     // 1 / (2 pi R T_w)^(D/2)
    double A[14];   // rho, rhoU(D), rhoE, Q(D), P(D(D+1)/2) of the incoming part
    double rho_b;   // open faces: boundary density
    double mode;    // open faces: 0 far field, 1 fixed (incoming U = U above; else the owner
                    // cell's), 2 pressure inlet, 3 pressure outlet (rho_b, T reset every step)
};
static_assert(sizeof(WallC) == 176, "wallc size");

struct EqC {
    double U[3];
    double invRT, pre, qc[3], RT;
};

// A discrete velocity is stored as its packed axis indices ix | iy << 8 | iz << 16 (bit 31 marks a
// padding slot, weight 0); xi and the weight come from per-axis tables staged in LDS.
constexpr uint32_t VPAD = 0x80000000u;

// Least-squares stencil of one cell, laid out for scalar loads: entry j is the cell's j-th slot
// (ascending face order).  The distribution buffers hold the cell rows followed by one row per
// wall face with the lagged fixedGradient gradient, and off[j] is the element offset of entry
// j's row: the other cell (internal, cyclic) or the wall row.  Entries past the cell's face count
// point at the cell's own row with L = 0 and contribute nothing.
struct alignas(64) CellSt {
    int64_t off[MAXS];
    double L[MAXS][3];
    double a[MAXS];       // wall entries: 1/deltaCoeffs
    int32_t wmask;        // bit j set: entry j is a wall face
    int32_t inner;        // 1: every slot is an internal face (KIND_OWN or KIND_NEI)
    int32_t pad[14];
};
static_assert(sizeof(CellSt) % 64 == 0, "cellst size");

// Per-face data of k_face.
struct alignas(16) FaceX {
    double S[3];
    double dP[3];   // Cf - C(owner)
    double dN[3];   // Cf - C(neighbour) | wall: unit normal Sf/|Sf|
    double w;       // cyclic: patch weight
    int32_t P, N;   // owner; neighbour or coupled partner cell; symmetry: mirror axis; else -1
    int32_t kind, bidx;  // KIND_OWN (internal), KIND_WALL, KIND_CYC; wall index
};
static_assert(sizeof(FaceX) == 96, "facex size");

template <typename TS>
struct P {
    int D, nc, nf, nvp, nxi, nax, K, nt, niter, nw;
    int nib, nb, vb;    // k_face velocity blocks: nb blocks of vb = nt * nib velocities
    int ntc, nibc, nbc, vbc;  // k_cell velocity blocks: nbc blocks of vbc = ntc * nibc velocities
    double R, Pr, omega, Tref, kc;  // kc = K + 3 - D
    const uint32_t *__restrict__ vidx;  // packed axis indices of each (sign-sorted) velocity
    const uint32_t *__restrict__ mir;   // [axis][velocity]: the velocity mirrored across that axis
    const double *__restrict__ axw;     // per-axis quadrature weights, same layout as axis
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
    const CellSt *__restrict__ cst;
    const FaceX *__restrict__ fx;
    const double *__restrict__ moff;  // FP32 storage: offsets [g: nvp | h: nvp] (see row_off)
    const int *__restrict__ corder;  // launch order of the cells
    const int *__restrict__ forder;  // launch order of the faces
    double *fpart;      // [face][block][moment] partial sums of k_face
    double *cpart;      // [cell][block][moment] partial sums of k_cell
    const TS *gb_in, *hb_in;
    TS *gb_out, *hb_out;
    const TS *gBg_in, *gBh_in;
    TS *gBg_out, *gBh_out;
    TS *fvg, *fvh;      // stored face values [face][velocity] (face-value mode), else null
    int nopen;          // symmetry, far-field, fixed and pressure faces
    const double *__restrict__ bpress;  // [wall record] pressureIn / pressureOut pressure
    const int *__restrict__ open;  // their face indices
    const int *__restrict__ fsel;  // k_face_moments over a face list (the wall faces), else null
    int nfsel;
    const int *__restrict__ mxf;   // k_face_moments_mx: non-wall faces in launch order, 16 per tile (-1 pads)
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

constexpr int NXCD = 8;  // accelerator complex dies of an MI355X, each with its own L2

template <typename TS>
DEVI double ld(const TS *a, size_t i)
{
    return (double)a[i];
}

// Store of a value that no later work of the same kernel reads (face values, the next step's
// gbar): a streaming store, so that it does not displace rows still being re-read from L2.
#ifndef DUGKS_NT_STORE
#define DUGKS_NT_STORE 1
#endif
// Load of a value that no other workgroup of the kernel reads.
#ifndef DUGKS_NT_LOAD
#define DUGKS_NT_LOAD 1
#endif
template <typename TS>
DEVI double ld_stream(const TS *a, size_t i)
{
    if (DUGKS_NT_LOAD) return (double)__builtin_nontemporal_load(a + i);
    return (double)a[i];
}

template <typename TS>
DEVI void st_stream(TS *a, size_t i, double v)
{
    if (DUGKS_NT_STORE) __builtin_nontemporal_store((TS)v, a + i);
    else a[i] = (TS)v;
}

struct Vel {
    double x, y, z, w, xi2;
    int ix, iy, iz;
};

// FP32 storage keeps cell rows as the deviation from a fixed rest Maxwellian M (density and
// temperature: volume means of the initial state): value = stored + M(xi).  FP32 rounding then
// acts on the small deviation only.  Least-squares gradients and wall rows are unaffected (M is
// a constant per velocity), so they work on the stored deviations directly.  M is the product
// of per-axis factors staged in LDS (vt[2 nax, 3 nax)) times a prefactor for g / h
// (vt[3 nax + which]).  FP64 storage keeps the values themselves (offset 0).
template <int D, typename TS>
DEVI double row_off(const P<TS> &p, const Vel &v, int which, const double *vt)
{
    if (sizeof(TS) == 8) return 0.0;
    const double *e = vt + 2 * p.nax;
    double m = e[v.ix];
    if (D >= 2) m *= e[p.nxi + v.iy];
    if (D == 3) m *= e[2 * p.nxi + v.iz];
    return m * vt[3 * p.nax + which];
}

// Per-axis velocity table in LDS: vt[0, nax) = axis values, vt[nax, 2 nax) = axis weights,
// vt[2 nax, 3 nax + 2) = the FP32 row-offset factors (row_off).
constexpr int VT_SIZE = 3 * NAXMAX + 2;

template <typename TS>
DEVI void stage_vt(const P<TS> &p, double *vt)
{
    for (int t = threadIdx.x; t < 3 * p.nax + 2; t += blockDim.x)
        vt[t] = t < p.nax ? p.axis[t] : (t < 2 * p.nax ? p.axw[t - p.nax] : p.moff[t - 2 * p.nax]);
}

// Velocity k: w = (w_z w_y) w_x (2-D: (w_y w_x) 1, 1-D: (w_x 1) 1) and |xi|^2 = (x x + y y) + z z in
// the order the velocity set is built (fvDVM.C:140-221), without contraction.
template <int D, typename TS>
DEVI Vel load_vel(const P<TS> &p, unsigned k, const double *vt)
{
#pragma clang fp contract(off)
    const uint32_t id = p.vidx[k];
    const int nxi = p.nxi, nax = p.nax;
    Vel v;
    v.ix = id & 0xff;
    v.iy = (id >> 8) & 0xff;
    v.iz = (id >> 16) & 0xff;
    v.x = vt[v.ix];
    v.y = D >= 2 ? vt[nxi + v.iy] : 0.0;
    v.z = D == 3 ? vt[2 * nxi + v.iz] : 0.0;
    const double wx = vt[nax + v.ix], wy = D >= 2 ? vt[nax + nxi + v.iy] : 1.0;
    v.w = D == 3 ? (vt[nax + 2 * nxi + v.iz] * wy) * wx : (D == 2 ? (wy * wx) * 1.0 : (wx * 1.0) * 1.0);
    if (id & VPAD) v.w = 0.0;
    v.xi2 = (v.x * v.x + v.y * v.y) + v.z * v.z;
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
    e.pre = (D == 3) ? rho / (tp * sqrt(tp)) : (D == 2 ? rho / tp : rho / sqrt(tp));
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
#pragma clang fp contract(on)
    double cx = e.U[0] - v.x, cy = e.U[1] - v.y, cz = e.U[2] - v.z;
    double c2 = (cx * cx + cy * cy + cz * cz) * e.invRT;
    double pc = -(cx * e.qc[0] + cy * e.qc[1] + cz * e.qc[2]);
    double gB = e.pre * tab[v.ix];
    if (D >= 2) gB *= tab[nxi + v.iy];
    if (D == 3) gB *= tab[2 * nxi + v.iz];
    geq = (1.0 + pc * ((c2 - D) - 2.0)) * gB;
    if (H) heq = (kc + pc * ((c2 - D) * kc - 2.0 * K)) * gB * e.RT;
    else heq = 0.0;
}

// S7 relaxation of one face value towards the face's Shakhov equilibrium: (1 - r) gf + r geq_f
// (discreteVelocity.C:867-920, 1016-1044).  tf points at the face's entry of the face-interleaved
// exp tables (entry a of the face at tf[a * 2 NS]); the x entries carry r pre, so the table product
// is r times the Gaussian part.  The polynomial factors use the face's expanded constants.
template <int D, bool H, typename FE>
DEVI void face_relax(const FE &fe, const double *tf, int xo, int yo, int zo, const Vel &v, double kc, int K,
                     double &gf, double &hf)
{
#pragma clang fp contract(on)
    double rgB = tf[xo];
    if (D >= 2) rgB *= tf[yo];
    if (D == 3) rgB *= tf[zo];
    const double P = v.xi2 * fe.invRT + fe.C0 + fe.B[0] * v.x + fe.B[1] * v.y + fe.B[2] * v.z;
    const double pc = fe.Q0 + fe.qc[0] * v.x + fe.qc[1] * v.y + fe.qc[2] * v.z;
    gf = fe.omr * gf + (1.0 + pc * P) * rgB;
    if (H) hf = fe.omr * hf + (kc + pc * ((P + 2.0) * kc - 2.0 * K)) * rgB * fe.RT;
}

// Mesh constants are read through the constant address space: with a wave-uniform index the
// loads become scalar loads whose results live in SGPRs, and they are never re-fetched per lane.
// Only arrays that no kernel of the launch writes are read this way.
#define CONST_AS __attribute__((address_space(4)))

template <typename T>
DEVI const CONST_AS T *cptr(const T *p)
{
    return (const CONST_AS T *)p;
}

// Least-squares gradient of gbar (and hbar) at one cell for velocity k
// (leastSquaresGrad: owner += pVec (phi_N - phi_P), neighbour -= nVec (phi_N - phi_P),
// non-coupled boundary += pVec_b (phi_b - phi_P) with phi_b = phi_P + gradient_b / deltaCoeffs_b).
// All NS stencil rows are loaded before any is used, so their latencies overlap; the stencil
// record comes in through scalar loads.  cell must be wave-uniform.
template <int D, bool H, int NS, typename TS>
DEVI void grad_at(const P<TS> &p, int cell, unsigned k, double pg, double ph, double *G, double *Hg)
{
    const CONST_AS CellSt &st = cptr(p.cst)[cell];
    double vg[NS], vh[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        vg[j] = ld(p.gb_in + st.off[j], k);
        vh[j] = H ? ld(p.hb_in + st.off[j], k) : 0.0;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        G[d] = 0.0;
        Hg[d] = 0.0;
    }
    const int wm = st.wmask;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double dg, dh = 0.0;
        if (wm & (1 << j)) {
            const double a = st.a[j];
            dg = (pg + vg[j] * a) - pg;
            if (H) dh = (ph + vh[j] * a) - ph;
        } else {
            dg = vg[j] - pg;
            if (H) dh = vh[j] - ph;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            G[d] = fma(st.L[j][d], dg, G[d]);
            if (H) Hg[d] = fma(st.L[j][d], dh, Hg[d]);
        }
    }
}

// a / b for a per-cell constant b with y = 1/b (rounded): q = a y, then one Newton correction with
// the exact residual a - b q (an FMA) gives the correctly rounded quotient (Markstein, IBM J. Res.
// Dev. 34 (1990) 111-119), so the S8 flux term s g dt / V rounds as the reference's division does.
DEVI double div_rn(double a, double b, double y)
{
    const double q = a * y;
    const double r = fma(-q, b, a);
    return fma(r, y, q);
}

// Shared scalar formulas.  "contract(on)" fuses multiply-adds only within one expression, so
// every kernel that uses them rounds identically (the default contraction is decided after
// inlining and may differ from kernel to kernel).
// S1 relaxation factor r = 1.5 dt / (2 tau + dt) (discreteVelocity.C:393)
DEVI double s1_r(double dt, double tau)
{
#pragma clang fp contract(on)
    return 1.5 * dt / (2.0 * tau + dt);
}

// heat-flux factor 2 tau / (2 tau + c) (fvDVM.C:522, 727)
DEVI double q_fac(double tau, double c)
{
#pragma clang fp contract(on)
    return 2.0 * tau / (2.0 * tau + c);
}

// S1 value (1 - r) g~ + r geq for a stored row value gv with row offset og (FP32 deviations)
DEVI double s1_value(double gv, double og, double r, double geq)
{
#pragma clang fp contract(on)
    return (1.0 - r) * (gv + og) + r * geq - og;
}

template <int D>
DEVI double dotg(const double *G, const double *dC, const double *hx)
{
    double r = G[0] * (dC[0] - hx[0]);
    if (D >= 2) r += G[1] * (dC[1] - hx[1]);
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
    const double hwg = hw * g;
    double t[3];
#pragma unroll
    for (int d = 0; d < D; ++d) t[d] = hwg * ch[d];
    if (H) {
        const double A = cc * g + h;
#pragma unroll
        for (int d = 0; d < D; ++d) a[D + 2 + d] = fma(hw * ch[d], A, a[D + 2 + d]);
    } else {
        // Q_d = sum (w g / 2) c_d |c|^2 with the shared products t_d = (w g / 2) c_d
#pragma unroll
        for (int d = 0; d < D; ++d) a[D + 2 + d] = fma(t[d], cc, a[D + 2 + d]);
    }
    int m = 2 * D + 2;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = i; j < D; ++j) {
            a[m] = fma(t[i], ch[j], a[m]);
            ++m;
        }
}

// Macroscopic state from the reduced moments (fvDVM.C:493-522 / 694-727).
// Returns rho, U, T, tau, q (q before the bar-to-original correction factor).
template <int D>
DEVI void finish_moments(const double *a, const double *Uh, double R, int K, double mu, double omega,
                         double Tref, double &rho, double *U, double &T, double &tau, double *q)
{
#pragma clang fp contract(on)
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
// Upwind face value (discreteVelocity.C:472-530 internal faces, 605-627 Maxwell wall outgoing,
// 628-672 coupled faces) from the gbar values, gradients and centre-to-face vectors of the face
// owner P and of the other cell N.  hx = xi dt / 2.
//   internal: xi.S >= VSMALL: P;  xi.S < -VSMALL: N;  otherwise the mean of both reconstructions
//   cyclic:   xi.S > VSMALL: P;   otherwise the coupled boundary value w phi_P + (1 - w) phi_N with
//             the coupled gradient, at distance -hx (the mean of both when |xi.S| <= VSMALL)
//   wall:     P (only used for xi.S > 0)
// ---------------------------------------------------------------------------------------------
template <int D>
DEVI double face_value(int kind, double s, double gP, const double *GP, const double *dP, double gN,
                       const double *GN, const double *dN, double w, const double *hx)
{
    if (kind == KIND_WALL) return gP + dotg<D>(GP, dP, hx);
    if (kind == KIND_CYC) {
        if (s > VSMALL) return gP + dotg<D>(GP, dP, hx);
        double Gb[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) Gb[d] = w * GP[d] + (1.0 - w) * GN[d];
        const double z3[3] = {0, 0, 0};
        const double gb = w * gP + (1.0 - w) * gN;
        const double dog = dotg<D>(Gb, z3, hx);
        if (s < -VSMALL) return gb + dog;
        return 0.5 * (((gP + dotg<D>(GP, dP, hx)) + gb) + dog);
    }
    if (s >= VSMALL) return gP + dotg<D>(GP, dP, hx);
    if (s < -VSMALL) return gN + dotg<D>(GN, dN, hx);
    return 0.5 * (((gN + dotg<D>(GN, dN, hx)) + gP) + dotg<D>(GP, dP, hx));
}

// Lanes of a face that need the gradient of the owner P / of the other cell N.
DEVI bool need_P(int kind, double s) { return kind != KIND_OWN || !(s < -VSMALL); }
DEVI bool need_N(int kind, double s)
{
    if (grad_row(kind)) return false;
    if (kind == KIND_CYC) return !(s > VSMALL);
    return !(s >= VSMALL);
}

// ---------------------------------------------------------------------------------------------
// Work decomposition of the step kernels: the velocity space is cut into nb blocks of vb = nt*nib
// consecutive (sign-sorted) velocities, and a workgroup handles one (face or cell, block) pair.
// The grid is 2-D: x runs over NXCD * per entity slots, y over the velocity blocks (a 1-D grid of
// entities x blocks can exceed the 32-bit work-item count of one dispatch).  Workgroups are
// dealt to the XCDs round-robin in linear id order, and NXCD divides the x extent, so XCD
// x = blockIdx.x mod NXCD walks entity positions [x*per, (x+1)*per) of the launch order for
// block 0, then block 1, ...  The launch order is a
// space-filling-curve order of the cell (face) centres, so the workgroups resident on one XCD at a
// time cover a compact patch of the mesh and one velocity block: the neighbour rows that the
// least-squares gradients read are then shared through that XCD's L2.  Returns false for the
// padding workgroups.
// ---------------------------------------------------------------------------------------------
DEVI bool xcd_block(int n, int nb, int &pos, int &b)
{
    const int per = (n + NXCD - 1) / NXCD;
    b = (int)blockIdx.y;
    pos = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
    return pos < n && b < nb;
}

// ---------------------------------------------------------------------------------------------
// k_face: S2 reconstruction and the S6 face-moment sums over one velocity block
// (discreteVelocity.C:412-627, fvDVM.C:456-516).  The face value of each velocity is rebuilt from
// the resident gbar of the upwind cell and its least-squares gradient (never stored); the moments
// about the previous face velocity (wall velocity on walls) and the wall outGoing sum are reduced
// across the lanes and written as the block's partial sums.  Wall faces also write the next
// step's fixedGradient boundary gradient n . grad(gbar) (discreteVelocity.C:436-470).
// ---------------------------------------------------------------------------------------------
template <int D, bool H, int NS, typename TS, bool FV>
__global__ __launch_bounds__(MAXNT) __attribute__((amdgpu_waves_per_eu(sizeof(TS) == 8 ? FACE_WAVES : FACE_WAVES_FP32, 8))) void k_face(P<TS> p)
{
    constexpr int NM = NMom<D>::v + 1;  // + wall outGoing
    __shared__ double red[MAXWAVES * NM];
    __shared__ double tot[NM];
    __shared__ double vt[VT_SIZE];
    int pos, b;
    if (!xcd_block(p.nf, p.nb, pos, b)) return;
    stage_vt(p, vt);
    __syncthreads();
    const int f = cptr(p.forder)[pos];
    const CONST_AS FaceX &fx = cptr(p.fx)[f];
    const int kind = fx.kind;
    const bool wall = kind == KIND_WALL;
    const int cP = fx.P, cN = fx.N;
    const int cas = cptr(p.ccase)[cP];
    const double dt = cptr(p.dtc)[cas];
    const double S[3] = {fx.S[0], fx.S[1], fx.S[2]};
    const double dP[3] = {fx.dP[0], fx.dP[1], fx.dP[2]};
    const double dN[3] = {fx.dN[0], fx.dN[1], fx.dN[2]};
    const double wcy = fx.w;
    double Uh[3];
    if (wall) {
        const CONST_AS WallC &W = cptr(p.wall)[fx.bidx];
        Uh[0] = W.U[0]; Uh[1] = W.U[1]; Uh[2] = W.U[2];
    } else {
        Uh[0] = cptr(p.fU)[3 * f]; Uh[1] = cptr(p.fU)[3 * f + 1]; Uh[2] = cptr(p.fU)[3 * f + 2];
    }
    const TS *__restrict__ rowP = p.gb_in + (size_t)cP * p.nvp;
    const TS *__restrict__ rowN = p.gb_in + (size_t)(cN >= 0 ? cN : cP) * p.nvp;
    const TS *__restrict__ hrowP = H ? p.hb_in + (size_t)cP * p.nvp : nullptr;
    const TS *__restrict__ hrowN = H ? p.hb_in + (size_t)(cN >= 0 ? cN : cP) * p.nvp : nullptr;
    double acc[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) acc[m] = 0.0;
    const unsigned k0 = (unsigned)b * p.vb + threadIdx.x;
    for (int it = 0; it < p.nib; ++it) {
        const unsigned k = k0 + it * p.nt;
        const Vel v = load_vel<D>(p, k, vt);
        const double s = dotnc(v.x, v.y, v.z, S);
        const double hx[3] = {0.5 * v.x * dt, 0.5 * v.y * dt, 0.5 * v.z * dt};
        double gP = 0.0, hP = 0.0, gN = 0.0, hN = 0.0;
        double GP[3] = {0, 0, 0}, HP[3] = {0, 0, 0}, GN[3] = {0, 0, 0}, HN[3] = {0, 0, 0};
        if (__any(need_P(kind, s))) {
            gP = ld(rowP, k);
            if (H) hP = ld(hrowP, k);
            grad_at<D, H, NS>(p, cP, k, gP, hP, GP, HP);
        }
        if (__any(need_N(kind, s))) {
            gN = ld(rowN, k);
            if (H) hN = ld(hrowN, k);
            grad_at<D, H, NS>(p, cN, k, gN, hN, GN, HN);
        }
        if (wall) {
            const size_t bi = (size_t)fx.bidx * p.nvp + k;
            double ng = GP[0] * dN[0] + GP[1] * dN[1];
            if (D == 3) ng += GP[2] * dN[2];
            p.gBg_out[bi] = (TS)ng;
            if (H) {
                double nh = HP[0] * dN[0] + HP[1] * dN[1];
                if (D == 3) nh += HP[2] * dN[2];
                p.gBh_out[bi] = (TS)nh;
            }
        }
        if (!wall || s > 0) {
            const double gfd = face_value<D>(kind, s, gP, GP, dP, gN, GN, dN, wcy, hx);
            const double hfd = H ? face_value<D>(kind, s, hP, HP, dP, hN, HN, dN, wcy, hx) : 0.0;
            if (FV) {
                p.fvg[(size_t)f * p.nvp + k] = (TS)gfd;
                if (H) p.fvh[(size_t)f * p.nvp + k] = (TS)hfd;
            }
            const double gf = gfd + row_off<D>(p, v, 0, vt);
            const double hf = H ? hfd + row_off<D>(p, v, 1, vt) : 0.0;
            acc_moments<D, H>(acc, gf, hf, v, Uh);
            if (wall) acc[NM - 1] = fma(v.w * s, gf, acc[NM - 1]);
        }
    }
    block_sum<NM>(acc, red, tot);
    if ((int)threadIdx.x < NM) p.fpart[((size_t)f * p.nb + b) * NM + threadIdx.x] = tot[threadIdx.x];
}

// Wave target of k_face_values (KFV_WAVES where no instantiation of that shape spills).
template <int D, bool H, int NS, typename TS>
constexpr int kfv_waves()
{
    return (!H && NS <= 8) || (D <= 2 && NS <= 4) ? KFV_WAVES : sizeof(TS) == 8 ? FACE_WAVES : FACE_WAVES_FP32;
}

// ---------------------------------------------------------------------------------------------
// Face-value mode, S2 in two passes.
// k_face_values: cell-wise, over one velocity block.  The least-squares gradient of the cell is
// formed once per velocity, and the cell writes the upwind face value of every face it is upwind
// of: internal faces it owns with xi.S >= -VSMALL (|xi.S| <= VSMALL is the mean of both
// reconstructions, for which the neighbour's gradient is formed on demand), internal faces it
// neighbours with xi.S < -VSMALL, its cyclic faces (the coupled value needs the partner cell's
// gradient where xi.S <= VSMALL) and the outgoing half of its wall faces, whose lagged
// fixedGradient rows n . grad(gbar) it writes as well (so also for symmetry, far-field and fixed
// faces, whose incoming halves k_face_open writes).  Every face value is written exactly once,
// with the arithmetic of k_face (same gradient, same face_value), so the stored values equal
// k_face's.  A cell whose faces are all internal (CellSt::inner) takes a copy of the internal
// branches unrolled over the stencil, without the boundary-kind dispatch of the general loop.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, int NS, typename TS>
__global__ __launch_bounds__(MAXNT) __attribute__((amdgpu_waves_per_eu(kfv_waves<D, H, NS, TS>(), 8))) void k_face_values(P<TS> p)
{
    __shared__ double vt[VT_SIZE];
    int pos, b;
    if (!xcd_block(p.nc, p.nb, pos, b)) return;
    stage_vt(p, vt);
    __syncthreads();
    const int c = cptr(p.corder)[pos];
    const int cas = cptr(p.ccase)[c];
    const double dt = cptr(p.dtc)[cas];
    const int s0 = cptr(p.cs)[c];
    const int ns = cptr(p.cs)[c + 1] - s0;
    const CONST_AS Slot *SL = cptr(p.slots) + s0;
    const TS *__restrict__ rowC = p.gb_in + (size_t)c * p.nvp;
    const TS *__restrict__ hrowC = H ? p.hb_in + (size_t)c * p.nvp : nullptr;
    const unsigned k0 = (unsigned)b * p.vb + threadIdx.x;
    if (KFV_INNER && cptr(p.cst)[c].inner) {
        // every face of the cell is internal: the owner and neighbour branches of the loop below,
        // unrolled over the stencil, with the same arithmetic
        for (int it = 0; it < p.nib; ++it) {
            const unsigned k = k0 + it * p.nt;
            const Vel v = load_vel<D>(p, k, vt);
            const double hx[3] = {0.5 * v.x * dt, 0.5 * v.y * dt, 0.5 * v.z * dt};
            const double gC = ld(rowC, k);
            const double hC = H ? ld(hrowC, k) : 0.0;
            double GC[3] = {0, 0, 0}, HC[3] = {0, 0, 0};
            grad_at<D, H, NS>(p, c, k, gC, hC, GC, HC);
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (j >= ns) continue;   // a constant trip count keeps the loop unrollable
                const CONST_AS Slot &S = SL[j];
                const double Sv[3] = {S.S[0], S.S[1], S.S[2]};
                const double s = dotnc(v.x, v.y, v.z, Sv);
                const double dCs[3] = {S.dCs[0], S.dCs[1], S.dCs[2]};
                TS *__restrict__ fg = p.fvg + (size_t)S.face * p.nvp;
                TS *__restrict__ fh = H ? p.fvh + (size_t)S.face * p.nvp : nullptr;
                if (S.kind == KIND_NEI) {
                    if (s < -VSMALL) {
                        st_stream(fg, k, gC + dotg<D>(GC, dCs, hx));
                        if (H) st_stream(fh, k, hC + dotg<D>(HC, dCs, hx));
                    }
                    continue;
                }
                // owner: upwind where xi.S >= VSMALL; where |xi.S| < VSMALL the mean of both
                // reconstructions, which needs the neighbour's gradient
                const bool mid = !(s >= VSMALL) && !(s < -VSMALL);
                if (__any(mid)) {
                    const double dCo[3] = {S.dCo[0], S.dCo[1], S.dCo[2]};
                    const int o = S.other;
                    const double gO = ld(p.gb_in + (size_t)o * p.nvp, k);
                    const double hO = H ? ld(p.hb_in + (size_t)o * p.nvp, k) : 0.0;
                    double GO[3] = {0, 0, 0}, HO[3] = {0, 0, 0};
                    grad_at<D, H, NS>(p, o, k, gO, hO, GO, HO);
                    if (mid) {
                        st_stream(fg, k, face_value<D>(KIND_OWN, s, gC, GC, dCs, gO, GO, dCo, 0.0, hx));
                        if (H) st_stream(fh, k, face_value<D>(KIND_OWN, s, hC, HC, dCs, hO, HO, dCo, 0.0, hx));
                    }
                }
                if (s >= VSMALL) {
                    st_stream(fg, k, gC + dotg<D>(GC, dCs, hx));
                    if (H) st_stream(fh, k, hC + dotg<D>(HC, dCs, hx));
                }
            }
        }
        return;
    }
    for (int it = 0; it < p.nib; ++it) {
        const unsigned k = k0 + it * p.nt;
        const Vel v = load_vel<D>(p, k, vt);
        const double hx[3] = {0.5 * v.x * dt, 0.5 * v.y * dt, 0.5 * v.z * dt};
        const double gC = ld(rowC, k);
        const double hC = H ? ld(hrowC, k) : 0.0;
        double GC[3] = {0, 0, 0}, HC[3] = {0, 0, 0};
        grad_at<D, H, NS>(p, c, k, gC, hC, GC, HC);
        for (int j = 0; j < ns; ++j) {
            const CONST_AS Slot &S = SL[j];
            const int kind = S.kind;
            const double Sv[3] = {S.S[0], S.S[1], S.S[2]};
            const double s = dotnc(v.x, v.y, v.z, Sv);
            const double dCs[3] = {S.dCs[0], S.dCs[1], S.dCs[2]};
            const double dCo[3] = {S.dCo[0], S.dCo[1], S.dCo[2]};
            TS *__restrict__ fg = p.fvg + (size_t)S.face * p.nvp;
            TS *__restrict__ fh = H ? p.fvh + (size_t)S.face * p.nvp : nullptr;
            if (kind == KIND_SPL) {
                // outgoing values reconstructed from this cell (discreteVelocity.C:673-689); the
                // incoming half (the mirrored velocities' values) is written by k_face_open
                if (s > 0) {
                    st_stream(fg, k, face_value<D>(KIND_WALL, s, gC, GC, dCs, 0.0, GC, dCo, 0.0, hx));
                    if (H) st_stream(fh, k, face_value<D>(KIND_WALL, s, hC, HC, dCs, 0.0, HC, dCo, 0.0, hx));
                }
                continue;
            }
            if (grad_row(kind)) {
                const size_t bi = (size_t)S.bidx * p.nvp + k;
                double ng = GC[0] * dCo[0] + GC[1] * dCo[1];
                if (D == 3) ng += GC[2] * dCo[2];
                p.gBg_out[bi] = (TS)ng;
                if (H) {
                    double nh = HC[0] * dCo[0] + HC[1] * dCo[1];
                    if (D == 3) nh += HC[2] * dCo[2];
                    p.gBh_out[bi] = (TS)nh;
                }
                if (kind == KIND_ZG) {
                    // zeroGradient: the cell's gBarPvol for every velocity (discreteVelocity.C:551-555)
                    st_stream(fg, k, gC);
                    if (H) st_stream(fh, k, hC);
                } else if (s > 0) {
                    st_stream(fg, k, face_value<D>(KIND_WALL, s, gC, GC, dCs, 0.0, GC, dCo, 0.0, hx));
                    if (H) st_stream(fh, k, face_value<D>(KIND_WALL, s, hC, HC, dCs, 0.0, HC, dCo, 0.0, hx));
                }
                continue;
            }
            if (kind == KIND_NEI) {
                // this cell is the face's neighbour N: upwind where xi.S < -VSMALL
                if (s < -VSMALL) {
                    const double z3[3] = {0, 0, 0};
                    st_stream(fg, k, face_value<D>(KIND_OWN, s, 0.0, z3, dCo, gC, GC, dCs, 0.0, hx));
                    if (H) st_stream(fh, k, face_value<D>(KIND_OWN, s, 0.0, z3, dCo, hC, HC, dCs, 0.0, hx));
                }
                continue;
            }
            // owner of an internal face or of a cyclic face
            const bool mine = kind == KIND_CYC || !(s < -VSMALL);
            const bool needO = kind == KIND_CYC ? !(s > VSMALL) : (mine && !(s >= VSMALL));
            double gO = 0.0, hO = 0.0;
            double GO[3] = {0, 0, 0}, HO[3] = {0, 0, 0};
            if (__any(needO)) {
                const int o = S.other;
                gO = ld(p.gb_in + (size_t)o * p.nvp, k);
                if (H) hO = ld(p.hb_in + (size_t)o * p.nvp, k);
                grad_at<D, H, NS>(p, o, k, gO, hO, GO, HO);
            }
            if (mine) {
                st_stream(fg, k, face_value<D>(kind, s, gC, GC, dCs, gO, GO, dCo, S.aux, hx));
                if (H) st_stream(fh, k, face_value<D>(kind, s, hC, HC, dCs, hO, HO, dCo, S.aux, hx));
            }
        }
    }
}

// k_face_open: the incoming values (xi.S <= 0) of symmetry, far-field and fixed faces, one
// workgroup per (face, velocity block).  Symmetry plane (DVMsymmetry, discreteVelocity.C:746-790;
// symmetryPlane, 791-815): the outgoing value of the velocity mirrored across the plane, rebuilt at
// its index from the owner cell's gbar and gradient.  Far field (discreteVelocity.C:588-602): rho_b M(U of the owner cell, T_b).
// Fixed (mixed, frozen at the initial boundary Maxwellian, discreteVelocity.C:312-344):
// rho_b M(U_b, T_b).  h = g R T_b (K + 3 - D) for the Maxwellians.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, int NS, typename TS>
__global__ __launch_bounds__(MAXNT) void k_face_open(P<TS> p)
{
    __shared__ double vt[VT_SIZE];
    const int i = blockIdx.x, b = blockIdx.y;
    if (i >= p.nopen || b >= p.nb) return;
    stage_vt(p, vt);
    __syncthreads();
    const int f = cptr(p.open)[i];
    const CONST_AS FaceX &fx = cptr(p.fx)[f];
    const int c = fx.P;
    const double dt = cptr(p.dtc)[cptr(p.ccase)[c]];
    const double S[3] = {fx.S[0], fx.S[1], fx.S[2]};
    const double dP[3] = {fx.dP[0], fx.dP[1], fx.dP[2]};
    const CONST_AS WallC &W = cptr(p.wall)[fx.bidx];
    const bool sym = fx.kind == KIND_SYM || fx.kind == KIND_SPL;
    const bool fixedU = W.mode == 1.0;
    const double Ui[3] = {fixedU ? W.U[0] : cptr(p.cU)[3 * c], fixedU ? W.U[1] : cptr(p.cU)[3 * c + 1],
                          fixedU ? W.U[2] : cptr(p.cU)[3 * c + 2]};
    const double RT = p.R * W.T;
    const TS *__restrict__ rowC = p.gb_in + (size_t)c * p.nvp;
    const TS *__restrict__ hrowC = H ? p.hb_in + (size_t)c * p.nvp : nullptr;
    TS *__restrict__ fg = p.fvg + (size_t)f * p.nvp;
    TS *__restrict__ fh = H ? p.fvh + (size_t)f * p.nvp : nullptr;
    const unsigned k0 = (unsigned)b * p.vb + threadIdx.x;
    for (int it = 0; it < p.nib; ++it) {
        const unsigned k = k0 + it * p.nt;
        const Vel v = load_vel<D>(p, k, vt);
        const double s = dotnc(v.x, v.y, v.z, S);
        if (s > 0) continue;
        if (sym) {
            const unsigned km = p.mir[(size_t)fx.N * p.nvp + k];
            const Vel vm = load_vel<D>(p, km, vt);
            const double hxm[3] = {0.5 * vm.x * dt, 0.5 * vm.y * dt, 0.5 * vm.z * dt};
            const double gm = ld(rowC, km);
            const double hm = H ? ld(hrowC, km) : 0.0;
            double GM[3] = {0, 0, 0}, HM[3] = {0, 0, 0};
            grad_at<D, H, NS>(p, c, km, gm, hm, GM, HM);
            st_stream(fg, k, gm + dotg<D>(GM, dP, hxm));
            if (H) st_stream(fh, k, hm + dotg<D>(HM, dP, hxm));
        } else {
            const double cx = Ui[0] - v.x, cy = Ui[1] - v.y, cz = Ui[2] - v.z;
            const double gin = W.rho_b * (W.pre * exp(-(cx * cx + cy * cy + cz * cz) / (2.0 * RT)));
            st_stream(fg, k, gin - row_off<D>(p, v, 0, vt));
            if (H) st_stream(fh, k, gin * RT * p.kc - row_off<D>(p, v, 1, vt));
        }
    }
}

// k_face_moments: face-wise, over one velocity block: the S6 moment sums and the wall outGoing
// sum of the stored face values (the moment part of k_face, same arithmetic and block split).
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_face_moments(P<TS> p)
{
    constexpr int NM = NMom<D>::v + 1;  // + wall outGoing
    __shared__ double red[MAXWAVES * NM];
    __shared__ double tot[NM];
    __shared__ double vt[VT_SIZE];
    int pos, b;
    if (!xcd_block(p.fsel ? p.nfsel : p.nf, p.nb, pos, b)) return;
    stage_vt(p, vt);
    __syncthreads();
    const int f = p.fsel ? cptr(p.fsel)[pos] : cptr(p.forder)[pos];
    const CONST_AS FaceX &fx = cptr(p.fx)[f];
    const bool wall = fx.kind == KIND_WALL;
    const double S[3] = {fx.S[0], fx.S[1], fx.S[2]};
    double Uh[3];
    if (wall) {
        const CONST_AS WallC &W = cptr(p.wall)[fx.bidx];
        Uh[0] = W.U[0]; Uh[1] = W.U[1]; Uh[2] = W.U[2];
    } else {
        Uh[0] = cptr(p.fU)[3 * f]; Uh[1] = cptr(p.fU)[3 * f + 1]; Uh[2] = cptr(p.fU)[3 * f + 2];
    }
    const TS *__restrict__ fg = p.fvg + (size_t)f * p.nvp;
    const TS *__restrict__ fh = H ? p.fvh + (size_t)f * p.nvp : nullptr;
    double acc[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) acc[m] = 0.0;
    const unsigned k0 = (unsigned)b * p.vb + threadIdx.x;
    // FM_UNR iterations per pass: their loads are issued before any is used; the moments are
    // accumulated in iteration order, as in k_face
    for (int it0 = 0; it0 < p.nib; it0 += FM_UNR) {
        double gv[FM_UNR], hv[FM_UNR];
#pragma unroll
        for (int u = 0; u < FM_UNR; ++u) {
            const unsigned k = k0 + (it0 + u < p.nib ? it0 + u : it0) * p.nt;
            gv[u] = ld(fg, k);
            hv[u] = H ? ld(fh, k) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < FM_UNR; ++u) {
            if (it0 + u >= p.nib) break;
            const unsigned k = k0 + (it0 + u) * p.nt;
            const Vel v = load_vel<D>(p, k, vt);
            const double s = dotnc(v.x, v.y, v.z, S);
            if (!wall || s > 0) {
                const double gf = gv[u] + row_off<D>(p, v, 0, vt);
                const double hf = H ? hv[u] + row_off<D>(p, v, 1, vt) : 0.0;
                acc_moments<D, H>(acc, gf, hf, v, Uh);
                if (wall) acc[NM - 1] = fma(v.w * s, gf, acc[NM - 1]);
            }
        }
    }
    block_sum<NM>(acc, red, tot);
    if ((int)threadIdx.x < NM) p.fpart[((size_t)f * p.nb + b) * NM + threadIdx.x] = tot[threadIdx.x];
}

// ---------------------------------------------------------------------------------------------
// k_face_moments_mx: the S6 moment sums of 16 non-wall faces per workgroup as FP64 matrix products
// (face-value mode).  The sums are a contraction over velocity space,
//   M[face][m] = sum_k fv[face][k] W[k][m],  W[k][m] = w_k phi_m(xi_k - Uh),
// with phi_m the moment weights of acc_moments (1, xi, |xi|^2 / 2, ch |ch|^2 / 2, ch ch^T / 2;
// h: 1 / 2, ch / 2) about one reference velocity Uh per tile: the previous velocity of the
// tile's first face (faces of a tile are consecutive in the Morton launch order, so their
// velocities are close), which is written to fU of each face of the tile for k_face_fin's exact
// shift to the new face velocity.
// Each wave sweeps 64-velocity chunks (wave w: chunks w, w + 4, ...): every lane computes the
// weights of one velocity into LDS and loads 16 consecutive face values of one face (a face's 64
// values per chunk are one coalesced 512 B segment), then 16 v_mfma_f64_16x16x4 accumulate
// D[face][m] += sum over 4 velocities of A[face][k] B[k][m] with A from registers and B from LDS.
// Operand layout of v_mfma_f64_16x16x4f64 (checked on gfx950): A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i mod 4), element i / 4.  The four wave
// partials are added in wave order.  FP32 storage keeps face values as deviations from the rest
// Maxwellian (row_off): a tile then holds 15 faces and row 15 carries the offsets, whose moments
// are added to every face's (sum (s + o) W = sum s W + sum o W).
// ---------------------------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int MX_NT = 256;
constexpr int MX_WS = 17;   // LDS row stride of the weight chunk (doubles), odd against bank conflicts

template <int D, bool H, typename TS>
__global__ __launch_bounds__(MX_NT) void k_face_moments_mx(P<TS> p)
{
    constexpr bool OFF = sizeof(TS) == 4;   // row 15 holds the FP32 row offsets
    constexpr int NMV = NMom<D>::v;
    constexpr int NM = NMV + 1;
    static_assert(NMV <= 16, "moments per MFMA column block");
    __shared__ double vt[VT_SIZE];
    __shared__ double Wl[MX_NT / 64][64 * MX_WS];
    __shared__ double Wh[H ? MX_NT / 64 : 1][H ? 64 * 4 : 1];
    // FP32 row offsets of the chunk: g in column 16 (the padding) of the weight rows, h here
    __shared__ double Ol[OFF && H ? MX_NT / 64 : 1][OFF && H ? 64 : 1];
    static_assert(64 * MX_WS >= 16 * 16, "the wave partials reuse the weight chunks");
    const int tile = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, kk = lane >> 4;
    stage_vt(p, vt);
    const int f0 = p.mxf[16 * tile];
    const int fi = p.mxf[16 * tile + i];
    const double Uh[3] = {p.fU[3 * f0], p.fU[3 * f0 + 1], p.fU[3 * f0 + 2]};
    __syncthreads();
    const TS *__restrict__ rowg = p.fvg + (size_t)(fi >= 0 ? fi : f0) * p.nvp;
    const TS *__restrict__ rowh = H ? p.fvh + (size_t)(fi >= 0 ? fi : f0) * p.nvp : nullptr;
    const double live = fi >= 0 ? 1.0 : 0.0;
    const bool offrow = OFF && i == 15;
    double *W = Wl[wave];
    v4d c = {0.0, 0.0, 0.0, 0.0};
    const int nchunk = p.nvp >> 6;
    for (int r = 0; 4 * r < nchunk; ++r) {
        const int chunk = 4 * r + wave;
        const bool act = chunk < nchunk;
        const int k0 = chunk << 6;
        double a[16], ah[16];
        if (act) {
            // weights of velocity k0 + lane
            const Vel v = load_vel<D>(p, (unsigned)(k0 + lane), vt);
            const double xi[3] = {v.x, v.y, v.z};
            double ch[3] = {0.0, 0.0, 0.0}, cc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ch[d] = xi[d] - Uh[d];
                cc = fma(ch[d], ch[d], cc);
            }
            const double hw = 0.5 * v.w;
            double *wr = W + lane * MX_WS;
            wr[0] = v.w;
#pragma unroll
            for (int d = 0; d < D; ++d) wr[1 + d] = v.w * xi[d];
            wr[D + 1] = hw * v.xi2;
#pragma unroll
            for (int d = 0; d < D; ++d) wr[D + 2 + d] = hw * ch[d] * cc;
            int m = 2 * D + 2;
#pragma unroll
            for (int a0 = 0; a0 < D; ++a0)
#pragma unroll
                for (int a1 = a0; a1 < D; ++a1) wr[m++] = hw * ch[a0] * ch[a1];
#pragma unroll
            for (int j = NMV; j < 16; ++j) wr[j] = 0.0;
            if (H) {
                double *wh = Wh[wave] + lane * 4;
                wh[0] = hw;
#pragma unroll
                for (int d = 0; d < 3; ++d) wh[1 + d] = d < D ? hw * ch[d] : 0.0;
            }
            if (OFF) {
                wr[16] = row_off<D>(p, v, 0, vt);
                if (H) Ol[wave][lane] = row_off<D>(p, v, 1, vt);
            }
            // face values k0 + 16 kk + m of face i (row 15 in FP32 storage: the row offsets,
            // read from LDS after the barrier)
            if (!offrow) {
                const TS *ga = rowg + k0 + 16 * kk;
#pragma unroll
                for (int m2 = 0; m2 < 16; ++m2) a[m2] = live * (double)ga[m2];
                if (H) {
                    const TS *ha = rowh + k0 + 16 * kk;
#pragma unroll
                    for (int m2 = 0; m2 < 16; ++m2) ah[m2] = live * (double)ha[m2];
                }
            }
        }
        __syncthreads();
        if (act) {
            if (offrow) {
#pragma unroll
                for (int m2 = 0; m2 < 16; ++m2) {
                    a[m2] = W[(16 * kk + m2) * MX_WS + 16];
                    ah[m2] = H ? Ol[wave][16 * kk + m2] : 0.0;
                }
            }
            // B[kk][j] of product m2 = W[16 kk + m2][j], j = this lane's column i
            const double *wb = W + (16 * kk) * MX_WS + i;
#pragma unroll
            for (int m2 = 0; m2 < 16; ++m2)
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m2], wb[m2 * MX_WS], c, 0, 0, 0);
            if (H) {
                // h enters the energy (column D + 1) and Q (columns D + 2 .. 2 D + 1)
                const int sel = i == D + 1 ? 0 : (i >= D + 2 && i < 2 * D + 2 ? i - D - 1 : -1);
                const double *hb = Wh[wave] + (16 * kk) * 4 + (sel < 0 ? 0 : sel);
                const double on = sel < 0 ? 0.0 : 1.0;
#pragma unroll
                for (int m2 = 0; m2 < 16; ++m2)
                    c = __builtin_amdgcn_mfma_f64_16x16x4f64(ah[m2], on * hb[m2 * 4], c, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // the wave partials D[row][j] go to the weight chunks (free after the loop's last barrier)
    double(*Cr)[64 * MX_WS] = Wl;
#pragma unroll
    for (int r = 0; r < 4; ++r) Cr[wave][(kk + 4 * r) * 16 + i] = c[r];
    __syncthreads();
    {
        const int row = threadIdx.x >> 4, j = threadIdx.x & 15;
        const int f = p.mxf[16 * tile + row];
        if (f >= 0) {
            double sum = ((Cr[0][row * 16 + j] + Cr[1][row * 16 + j]) + Cr[2][row * 16 + j]) + Cr[3][row * 16 + j];
            if (OFF) sum += ((Cr[0][15 * 16 + j] + Cr[1][15 * 16 + j]) + Cr[2][15 * 16 + j]) + Cr[3][15 * 16 + j];
            double *fp = p.fpart + (size_t)f * p.nb * NM;
            if (j < NM) fp[j] = j < NMV ? sum : 0.0;
            for (int b = 1; b < p.nb; ++b)
                if (j < NM) fp[b * NM + j] = 0.0;
            if (j < 3) p.fU[3 * f + j] = Uh[j];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// k_face_fin: one thread per face.  Sums the block partials in block order, then S3 wall density
// rho_w = outGoing / |inComingByRho| (calculatedMaxwellFvPatchField.C:139-164) with the moments of
// the S4 incoming Maxwellian added, the S6 face macros (fvDVM.C:456-522), the S7 equilibrium
// constants and relaxation factor, and the wall diagnostics qWall, stressWall (fvDVM.C:539-581).
// ---------------------------------------------------------------------------------------------
template <int D, typename TS>
__global__ __launch_bounds__(256) void k_face_fin(P<TS> p)
{
    constexpr int NM = NMom<D>::v + 1;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= p.nf) return;
    const FaceX &fx = p.fx[f];
    const bool wall = fx.kind == KIND_WALL;
    const int cas = p.ccase[fx.P];
    const double dt = p.dtc[cas];
    double a[NM];
    for (int m = 0; m < NM; ++m) a[m] = 0.0;
    const double *fp = p.fpart + (size_t)f * p.nb * NM;
    for (int b = 0; b < p.nb; ++b)
        for (int m = 0; m < NM; ++m) a[m] += fp[b * NM + m];
    double Uh[3];
    double rho_w = 0.0;
    if (wall) {
        const WallC &W = p.wall[fx.bidx];
        Uh[0] = W.U[0]; Uh[1] = W.U[1]; Uh[2] = W.U[2];
        rho_w = a[NM - 1] / fabs(W.incoming);
        for (int m = 0; m < NM - 1; ++m) a[m] = fma(rho_w, W.A[m], a[m]);
    } else {
        Uh[0] = p.fU[3 * f]; Uh[1] = p.fU[3 * f + 1]; Uh[2] = p.fU[3 * f + 2];
    }
    double rho, U[3], T, tau, q[3];
    finish_moments<D>(a, Uh, p.R, p.K, p.muc[cas], p.omega, p.Tref, rho, U, T, tau, q);
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
    double uu = 0.0, uq = 0.0;
    for (int d = 0; d < 3; ++d) {
        fe.B[d] = -2.0 * e.U[d] * e.invRT;
        uu += e.U[d] * e.U[d];
        uq += e.U[d] * e.qc[d];
    }
    fe.C0 = uu * e.invRT - (D + 2);
    fe.Q0 = -uq;
    fe.omr = 1.0 - fe.r;
    fe.pad2[0] = fe.pad2[1] = 0.0;
    if (wall) {
        const WallC &W = p.wall[fx.bidx];
        fe.pad[0] = W.T;
        fe.pad[1] = W.pre;
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
        for (int d = 0; d < 3; ++d) p.qwall[3 * fx.bidx + d] = d < D ? fw * a[D + 2 + d] : 0.0;
        for (int i = 0; i < 3; ++i)
            for (int jj = 0; jj < 3; ++jj) {
                double sv = 2.0 * Pm[i][jj] + W.U[i] * rU[jj] + W.U[jj] * rU[i] - W.U[i] * W.U[jj] * rho;
                p.swall[9 * fx.bidx + 3 * i + jj] = fs * sv;
            }
    }
    p.feq[f] = fe;
}

// ---------------------------------------------------------------------------------------------
// k_cell: S7 face relaxation, S8 g~ update and the S9 cell-moment sums over one velocity block
// (discreteVelocity.C:819-978, fvDVM.C:604-622, 712-721).  For each face of the cell (ascending
// face order, the order of the reference's face loops) the upwind face value is rebuilt exactly as
// k_face built it, relaxed towards the face equilibrium (incoming wall values: rho_w times the
// wall Maxwellian, left unrelaxed, discreteVelocity.C:693-731) and its flux applied.  g~ of this
// step is recovered from the resident gbar as (gbar - r geq) / (1 - r); the new g~ goes to the
// output buffer, where k_cell_s1 turns it into the next gbar.  With F1 (face-value mode, one
// workgroup per cell) the kernel also does the work of k_cell_fin and k_cell_s1 for its cell.
// Dynamic LDS: [face tables (2 NS nax) | velocity table (3 nax + 2) | tabC (nax)].  The face tables
// come first, so their addresses are a lane offset plus compile-time constants.
// ---------------------------------------------------------------------------------------------
template <int D, bool H, int NS, typename TS, bool FV, bool F1 = false>
__global__ __launch_bounds__(MAXNT) __attribute__((amdgpu_waves_per_eu(FV ? CELL_WAVES_FV : CELL_WAVES, 8))) void k_cell(P<TS> p)
{
    constexpr int NC = NMom<D>::v;
    extern __shared__ double tabs[];
    __shared__ double red[MAXWAVES * NC];
    __shared__ double tot[NC];

    int pos, b;
    if (!xcd_block(p.nc, p.nbc, pos, b)) return;
    const int c = cptr(p.corder)[pos];
    const int nax = p.nax;
    // face-interleaved tables: relaxation entry a of face j at tabF[a 2NS + j], wall entry at
    // tabW[a 2NS + j]
    double *tabF = tabs;
    double *tabW = tabs + NS;
    double *vt = tabs + 2 * NS * nax;
    double *tabC = vt + 3 * nax + 2;
    stage_vt(p, vt);
    const int cas = cptr(p.ccase)[c];
    const double dt = cptr(p.dtc)[cas];
    const double Vc = cptr(p.cV)[c];
    const double invV = 1.0 / Vc;
    const int s0 = cptr(p.cs)[c];
    const int ns = cptr(p.cs)[c + 1] - s0;
    const CONST_AS Slot *SL = cptr(p.slots) + s0;
    const TS *__restrict__ rowC = p.gb_in + (size_t)c * p.nvp;
    const TS *__restrict__ hrowC = H ? p.hb_in + (size_t)c * p.nvp : nullptr;

    // state that formed gbar: macros of the previous step
    const double Uc[3] = {cptr(p.cU)[3 * c], cptr(p.cU)[3 * c + 1], cptr(p.cU)[3 * c + 2]};
    const double qcv[3] = {cptr(p.cq)[3 * c], cptr(p.cq)[3 * c + 1], cptr(p.cq)[3 * c + 2]};
    EqC ec;
    eq_consts(cptr(p.crho)[c], Uc, cptr(p.cT)[c], qcv, p.R, p.Pr, D, ec);
    const double rc = s1_r(dt, cptr(p.ctau)[c]);
    const double inv1mr = 1.0 / (1.0 - rc);
    for (int t = threadIdx.x; t < nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = Uc[ax] - p.axis[t - ax * p.nxi];
        tabC[t] = exp(-(u * u) * (0.5 * ec.invRT));
    }
    // face equilibrium tables of every face and the wall Maxwellian tables of the wall faces,
    // interleaved by face so that the velocity loop addresses
    // every face's entry from one lane offset; the x entries carry the factors of the whole
    // product (r pre for the relaxation, rho_w pre_w for the incoming wall Maxwellian)
    for (int t = threadIdx.x; t < ns * nax; t += blockDim.x) {
        const int j = t / nax, a = t - j * nax;
        const int ax = a / p.nxi;
        const Slot &sj = p.slots[s0 + j];
        const FaceEq &fe = p.feq[sj.face];
        const double u = fe.U[ax] - p.axis[a - ax * p.nxi];
        double e = exp(-(u * u) * (0.5 * fe.invRT));
        if (ax == 0) e *= fe.r * fe.pre;
        tabF[a * 2 * NS + j] = e;
        if (sj.kind == KIND_WALL) {
            double w = p.wexp[(size_t)sj.bidx * nax + a];
            if (ax == 0) w *= fe.rho_w * fe.pad[1];
            tabW[a * 2 * NS + j] = w;
        }
    }
    __syncthreads();

    double acc[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] = 0.0;
    const unsigned k0 = (unsigned)b * p.vbc + threadIdx.x;
    for (int it = 0; it < p.nibc; ++it) {
        const unsigned k = k0 + it * p.ntc;
        const Vel v = load_vel<D>(p, k, vt);
        const int xo = v.ix * 2 * NS, yo = (p.nxi + v.iy) * 2 * NS, zo = (2 * p.nxi + v.iz) * 2 * NS;
        const double hx[3] = {0.5 * v.x * dt, 0.5 * v.y * dt, 0.5 * v.z * dt};
        // stored value (FP32: deviation); in face-value mode no other workgroup reads this row
        const double gC = FV ? ld_stream(rowC, k) : ld(rowC, k);
        const double hC = H ? (FV ? ld_stream(hrowC, k) : ld(hrowC, k)) : 0.0;
        const double og = row_off<D>(p, v, 0, vt), oh = H ? row_off<D>(p, v, 1, vt) : 0.0;
        const double gCv = gC + og, hCv = hC + oh;      // gbar itself
        double GC[3] = {0, 0, 0}, HC[3] = {0, 0, 0};
        if (!FV) grad_at<D, H, NS>(p, c, k, gC, hC, GC, HC);
        // g~ of this step from gbar (inverse of S1)
        double geq, heq;
        eq_eval<D, H>(ec, tabC, p.nxi, v, p.kc, p.K, geq, heq);
        const double gt = (gCv - rc * geq) * inv1mr;
        const double ht = H ? (hCv - rc * heq) * inv1mr : 0.0;
        // S8 (discreteVelocity.C:937-975): -1/3 g~ + 4/3 gbar, then the faces in face order
        double ag = -1.0 / 3 * gt + 4.0 / 3 * gCv;
        double ah = H ? (-1.0 / 3 * ht + 4.0 / 3 * hCv) : 0.0;
        // face-value mode: the values of all faces are requested before the first is used
        constexpr bool PRE = FV && KC_PRELOAD;
        double pg[PRE ? NS : 1], pgh[PRE ? NS : 1];
        if (PRE) {
#pragma unroll
            for (int j = 0; j < (PRE ? NS : 1); ++j) {
                const int fj = SL[j < ns ? j : 0].face;
                pg[j] = ld(p.fvg + (size_t)fj * p.nvp, k);
                pgh[j] = H ? ld(p.fvh + (size_t)fj * p.nvp, k) : 0.0;
            }
        }
        constexpr int FUNR = PRE ? NS : 1;
#pragma unroll FUNR
        for (int j = 0; j < (PRE ? NS : ns); ++j) {
            if (PRE && j >= ns) break;
            const CONST_AS Slot &S = SL[j];
            const int kind = S.kind;
            const double Sv[3] = {S.S[0], S.S[1], S.S[2]};
            const double s = dotnc(v.x, v.y, v.z, Sv);
            // roles: this cell is the face owner P except on KIND_NEI faces
            const bool nei = kind == KIND_NEI;
            bool needO;
            if (grad_row(kind) || kind == KIND_SPL) needO = false;
            else if (nei) needO = need_P(KIND_OWN, s);
            else needO = need_N(kind, s);
            double gO = 0.0, hO = 0.0;
            double GO[3] = {0, 0, 0}, HO[3] = {0, 0, 0};
            if (!FV && __any(needO)) {
                const int o = S.other;
                gO = ld(p.gb_in + (size_t)o * p.nvp, k);
                if (H) hO = ld(p.hb_in + (size_t)o * p.nvp, k);
                grad_at<D, H, NS>(p, o, k, gO, hO, GO, HO);
            }
            const double dCs[3] = {S.dCs[0], S.dCs[1], S.dCs[2]};
            const double dCo[3] = {S.dCo[0], S.dCo[1], S.dCo[2]};
            const CONST_AS FaceEq &fe = cptr(p.feq)[S.face];
            double gf, hf = 0.0;
            if (kind == KIND_WALL && !(s > 0)) {
                // S4 incoming Maxwellian rho_w M_w, left unrelaxed by S7 (maxwellWall operators
                // are no-ops); the x entries of the table carry rho_w pre_w
                const double *tw = tabW + j;
                gf = tw[xo];
                if (D >= 2) gf *= tw[yo];
                if (D == 3) gf *= tw[zo];
                if (H) hf = gf * (p.R * fe.pad[0]) * p.kc;
            } else {
                if (PRE) {
                    gf = pg[j] + og;
                    if (H) hf = pgh[j] + oh;
                } else if (FV) {
                    // the stored face value (same arithmetic as the rebuilding branches)
                    gf = ld(p.fvg + (size_t)S.face * p.nvp, k) + og;
                    if (H) hf = ld(p.fvh + (size_t)S.face * p.nvp, k) + oh;
                } else if (nei) {
                    gf = face_value<D>(KIND_OWN, s, gO, GO, dCo, gC, GC, dCs, 0.0, hx) + og;
                    if (H) hf = face_value<D>(KIND_OWN, s, hO, HO, dCo, hC, HC, dCs, 0.0, hx) + oh;
                } else {
                    gf = face_value<D>(kind, s, gC, GC, dCs, gO, GO, dCo, S.aux, hx) + og;
                    if (H) hf = face_value<D>(kind, s, hC, HC, dCs, hO, HO, dCo, S.aux, hx) + oh;
                }
                // S7 (discreteVelocity.C:867-920): relax towards the face equilibrium; on symmetry,
                // far-field and fixed patches only the outgoing values (their operators are no-ops);
                // on a symmetryPlane every value (the field operation, 880-881) and the outgoing
                // ones once more (the loop over non-coupled patches, 905-920)
                if (kind == KIND_OWN || kind == KIND_NEI || kind == KIND_CYC || kind == KIND_SPL || s > 0)
                    face_relax<D, H>(fe, tabF + j, xo, yo, zo, v, p.kc, p.K, gf, hf);
                if (kind == KIND_SPL && s > 0) face_relax<D, H>(fe, tabF + j, xo, yo, zo, v, p.kc, p.K, gf, hf);
            }
            // ag -/+ tg: the sign is exact, so the scalar-signed FMA rounds as the subtraction
            const double sgn = nei ? 1.0 : -1.0;
            ag = fma(sgn, div_rn(s * gf * dt, Vc, invV), ag);
            if (H) ah = fma(sgn, div_rn(s * hf * dt, Vc, invV), ah);
        }
        p.gb_out[(size_t)c * p.nvp + k] = (TS)(ag - og);
        if (H) p.hb_out[(size_t)c * p.nvp + k] = (TS)(ah - oh);
        acc_moments<D, H>(acc, ag, ah, v, Uc);
    }
    block_sum<NC>(acc, red, tot);
    if (!F1) {
        if ((int)threadIdx.x < NC) p.cpart[((size_t)c * p.nbc + b) * NC + threadIdx.x] = tot[threadIdx.x];
        return;
    }
    // F1 (one workgroup holds all velocities of the cell, nbc = 1): the work of k_cell_fin and
    // k_cell_s1 for this cell, with their arithmetic, on the row this workgroup has just written.
    // Each lane re-reads only the velocities it stored itself, from L2.
    double a[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) a[m] = 0.0 + tot[m];
    double rho, U[3], T, tau, q[3];
    finish_moments<D>(a, Uc, p.R, p.K, p.muc[cas], p.omega, p.Tref, rho, U, T, tau, q);
    const double fqc = q_fac(tau, dt * p.Pr);
    for (int d = 0; d < 3; ++d) q[d] = fqc * q[d];
    if (threadIdx.x == 0) {
        p.crho[c] = rho;
        p.cT[c] = T;
        p.ctau[c] = tau;
        for (int d = 0; d < 3; ++d) {
            p.cU[3 * c + d] = U[d];
            p.cq[3 * c + d] = q[d];
        }
        if (!(rho > 0.0) || !(T > 0.0) || !isfinite(rho) || !isfinite(T) || !isfinite(tau)) p.flag[0] = 1;
    }
    EqC en;
    eq_consts(rho, U, T, q, p.R, p.Pr, D, en);
    // every lane is past the velocity loop (block_sum ends in a barrier): tabC can be rebuilt
    for (int t = threadIdx.x; t < nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = U[ax] - p.axis[t - ax * p.nxi];
        tabC[t] = exp(-(u * u) * (0.5 * en.invRT));
    }
    __syncthreads();
    const double rn = s1_r(dt, tau);
    TS *__restrict__ go = p.gb_out + (size_t)c * p.nvp;
    TS *__restrict__ ho = H ? p.hb_out + (size_t)c * p.nvp : nullptr;
    for (int it = 0; it < p.nibc; ++it) {
        const unsigned k = k0 + it * p.ntc;
        const Vel v = load_vel<D>(p, k, vt);
        const double gv = ld(go, k), hv = H ? ld(ho, k) : 0.0;
        double geq, heq;
        eq_eval<D, H>(en, tabC, p.nxi, v, p.kc, p.K, geq, heq);
        const double og = row_off<D>(p, v, 0, vt), oh = H ? row_off<D>(p, v, 1, vt) : 0.0;
        st_stream(go, k, s1_value(gv, og, rn, geq));
        if (H) st_stream(ho, k, s1_value(hv, oh, rn, heq));
    }
}

// ---------------------------------------------------------------------------------------------
// k_cell_fin: one thread per cell.  Sums the block partials in block order and forms the S9 cell
// macros (fvDVM.C:604-727): rho, U, T, tau and q with the 2 tau / (2 tau + dt Pr) factor.
// ---------------------------------------------------------------------------------------------
template <int D, typename TS>
__global__ __launch_bounds__(256) void k_cell_fin(P<TS> p)
{
    constexpr int NC = NMom<D>::v;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nc) return;
    const int cas = p.ccase[c];
    const double dt = p.dtc[cas];
    double a[NC];
    for (int m = 0; m < NC; ++m) a[m] = 0.0;
    const double *cp = p.cpart + (size_t)c * p.nbc * NC;
    for (int b = 0; b < p.nbc; ++b)
        for (int m = 0; m < NC; ++m) a[m] += cp[b * NC + m];
    const double Uc[3] = {p.cU[3 * c], p.cU[3 * c + 1], p.cU[3 * c + 2]};
    double rho, U[3], T, tau, q[3];
    finish_moments<D>(a, Uc, p.R, p.K, p.muc[cas], p.omega, p.Tref, rho, U, T, tau, q);
    const double fqc = q_fac(tau, dt * p.Pr);
    for (int d = 0; d < 3; ++d) q[d] = fqc * q[d];
    p.crho[c] = rho;
    p.cT[c] = T;
    p.ctau[c] = tau;
    for (int d = 0; d < 3; ++d) {
        p.cU[3 * c + d] = U[d];
        p.cq[3 * c + d] = q[d];
    }
    if (!(rho > 0.0) || !(T > 0.0) || !isfinite(rho) || !isfinite(T) || !isfinite(tau)) p.flag[0] = 1;
}

// ---------------------------------------------------------------------------------------------
// k_pressure_bc: S10, updatePressureInOutBC (fvDVM.C:730-806), one thread per open face, after
// every step and once at initialisation (the constructor's call, fvDVM.C:1071-1072).  Pressure
// inlet: rho_b = p / (R T_b).  Pressure outlet, from the boundary cell's new state (rho_i, T_i),
// a_i = sqrt(R T_i (K + 5) / (K + 3)): rho_b = rho_i + (p - rho_i R T_i) / a_i^2, T_b = p / (R rho_i).
// The next step's incoming Maxwellian uses them (k_face_open).
// ---------------------------------------------------------------------------------------------
template <int D, typename TS>
__global__ __launch_bounds__(256) void k_pressure_bc(P<TS> p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.nopen) return;
    const FaceX &fx = p.fx[p.open[i]];
    WallC &W = p.wall[fx.bidx];
    const double pr = p.bpress[fx.bidx];
    if (W.mode == 2.0) {
        W.rho_b = pr / p.R / W.T;
    } else if (W.mode == 3.0) {
        const int c = fx.P;
        const double Ti = p.cT[c], rhoi = p.crho[c];
        const double ai = sqrt(p.R * Ti * (p.K + 5) / (p.K + 3));
        W.rho_b = rhoi + (pr - rhoi * p.R * Ti) / ai / ai;
        W.T = pr / (p.R * rhoi);
        const double tp = 2.0 * PI * (p.R * W.T);
        W.pre = D == 3 ? 1.0 / (tp * sqrt(tp)) : (D == 2 ? 1.0 / tp : 1.0 / sqrt(tp));
    }
}

// ---------------------------------------------------------------------------------------------
// k_cell_s1: S1 of the next step (discreteVelocity.C:393-406): gbar = (1 - r) g~ + r geq with the
// new cell macros, in place in the output buffer.  One workgroup per cell over all velocities
// (a pure stream of the row; the exp table is built once per cell).
// ---------------------------------------------------------------------------------------------
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_cell_s1(P<TS> p)
{
    __shared__ double tabN[NAXMAX];
    __shared__ double vt[VT_SIZE];
    const int c = blockIdx.x;
    stage_vt(p, vt);
    const int cas = cptr(p.ccase)[c];
    const double dt = cptr(p.dtc)[cas];
    const double Un[3] = {cptr(p.cU)[3 * c], cptr(p.cU)[3 * c + 1], cptr(p.cU)[3 * c + 2]};
    const double qn[3] = {cptr(p.cq)[3 * c], cptr(p.cq)[3 * c + 1], cptr(p.cq)[3 * c + 2]};
    EqC en;
    eq_consts(cptr(p.crho)[c], Un, cptr(p.cT)[c], qn, p.R, p.Pr, D, en);
    for (int t = threadIdx.x; t < p.nax; t += blockDim.x) {
        const int ax = t / p.nxi;
        const double u = Un[ax] - p.axis[t - ax * p.nxi];
        tabN[t] = exp(-(u * u) * (0.5 * en.invRT));
    }
    __syncthreads();
    const double rn = s1_r(dt, cptr(p.ctau)[c]);
    TS *__restrict__ go = p.gb_out + (size_t)c * p.nvp;
    TS *__restrict__ ho = H ? p.hb_out + (size_t)c * p.nvp : nullptr;
    // S1_UNR velocities per thread in flight: all loads of a group are issued before its stores
    constexpr int U = S1_UNR;
    for (unsigned k0 = threadIdx.x; k0 < (unsigned)p.nvp; k0 += U * blockDim.x) {
        double gv[U], hv[U];
        Vel vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned k = k0 + u * blockDim.x;
            gv[u] = hv[u] = 0.0;
            if (k < (unsigned)p.nvp) {
                vv[u] = load_vel<D>(p, k, vt);
                gv[u] = ld(go, k);
                if (H) hv[u] = ld(ho, k);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned k = k0 + u * blockDim.x;
            if (k < (unsigned)p.nvp) {
                double geq, heq;
                eq_eval<D, H>(en, tabN, p.nxi, vv[u], p.kc, p.K, geq, heq);
                const double og = row_off<D>(p, vv[u], 0, vt), oh = H ? row_off<D>(p, vv[u], 1, vt) : 0.0;
                st_stream(go, k, s1_value(gv[u], og, rn, geq));
                if (H) st_stream(ho, k, s1_value(hv[u], oh, rn, heq));
            }
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
    __shared__ double vt[VT_SIZE];
    stage_vt(p, vt);
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
        const Vel v = load_vel<D>(p, k, vt);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        p.gb_out[ck] = (TS)(geq - row_off<D>(p, v, 0, vt));
        if (H) p.hb_out[ck] = (TS)(heq - row_off<D>(p, v, 1, vt));
    }
}

// Converts between g~ and gbar = (1-r) g~ + r geq of the cell macros.
// mode 0: g~ -> gbar with dt_new; mode 1: gbar(dt_old) -> gbar(dt_new); mode 2: gbar(dt_old) -> g~.
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_convert(P<TS> p, const double *dt_old, const double *dt_new, int mode)
{
    __shared__ double tab[NAXMAX];
    __shared__ double vt[VT_SIZE];
    stage_vt(p, vt);
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
    if (mode != 0) ro = s1_r(dt_old[cas], tau);
    if (mode != 2) rn = s1_r(dt_new[cas], tau);
    const double inv1mr = 1.0 / (1.0 - ro);
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel<D>(p, k, vt);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        const double og = row_off<D>(p, v, 0, vt), oh = H ? row_off<D>(p, v, 1, vt) : 0.0;
        double g = ld(p.gb_in, ck) + og, h = H ? ld(p.hb_in, ck) + oh : 0.0;
        if (mode != 0) {
            g = (g - ro * geq) * inv1mr;
            h = (h - ro * heq) * inv1mr;
        }
        if (mode != 2) {
            g = (1.0 - rn) * g + rn * geq;
            h = (1.0 - rn) * h + rn * heq;
        }
        p.gb_out[ck] = (TS)(g - og);
        if (H) p.hb_out[ck] = (TS)(h - oh);
    }
}

// Wall set-up: axis tables of the wall Maxwellian, inComingByRho (fvDVM.C:263-309) and the
// moments of the incoming part per unit rho_w.
template <int D, typename TS>
__global__ __launch_bounds__(256) void k_wall_init(P<TS> p, const int *wall_face)
{
    constexpr int NM = NMom<D>::v;
    __shared__ double tab[NAXMAX];
    __shared__ double vt[VT_SIZE];
    stage_vt(p, vt);
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
        const Vel v = load_vel<D>(p, k, vt);
        const double s = dotnc(v.x, v.y, v.z, Sf);
        double m = W.pre * tab[v.ix];
        if (D >= 2) m *= tab[p.nxi + v.iy];
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
// ref_nv > 0: rows of ref_nv values in the reference velocity order (k = (iz nY + iy) nX + ix),
// padding slots skipped; else rows of nvp values in the engine's sign-sorted order.
template <int D, bool H, typename TS>
__global__ __launch_bounds__(MAXNT) void k_extract(P<TS> p, const int *cells, const double *dt_old, int raw,
                                                  double *g_out, double *h_out, int ref_nv)
{
    __shared__ double tab[NAXMAX];
    __shared__ double vt[VT_SIZE];
    stage_vt(p, vt);
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
    const double ro = raw ? 0.0 : s1_r(dt_old[cas], p.ctau[c]);
    const double inv1mr = 1.0 / (1.0 - ro);
    for (int it = 0; it < p.niter; ++it) {
        const int k = threadIdx.x + it * p.nt;
        const Vel v = load_vel<D>(p, k, vt);
        double geq, heq;
        eq_eval<D, H>(e, tab, p.nxi, v, p.kc, p.K, geq, heq);
        const size_t ck = (size_t)c * p.nvp + k;
        double g = ld(p.gb_in, ck) + row_off<D>(p, v, 0, vt), h = H ? ld(p.hb_in, ck) + row_off<D>(p, v, 1, vt) : 0.0;
        if (!raw) {
            g = (g - ro * geq) * inv1mr;
            h = (h - ro * heq) * inv1mr;
        }
        size_t o = (size_t)blockIdx.x * p.nvp + k;
        if (ref_nv > 0) {
            const uint32_t id = p.vidx[k];
            if (id & VPAD) continue;
            const int ix = id & 0xff, iy = (id >> 8) & 0xff, iz = (id >> 16) & 0xff;
            o = (size_t)blockIdx.x * ref_nv + ((size_t)iz * p.nxi + iy) * p.nxi + ix;
        }
        g_out[o] = g;
        h_out[o] = H ? h : 0.0;
    }
}

}  // namespace

// Device scratch of dugks_get_df_noalloc: the cell list and the g, h rows of one call.  One copy per
// GPU, held until the process ends; every call on a GPU shares it, so calls that overlap on one
// GPU race.  Host code reaches it only through the address hipGetSymbolAddress gives for the
// current device (df_scratch_ptr); the array's name is not a device address on the host.
constexpr size_t DF_SCRATCH_DOUBLES = DUGKS_DF_NOALLOC_MAX_DOUBLES;
static __device__ __align__(256) double scratch[DF_SCRATCH_DOUBLES];

// =============================================================================================
// host side
// =============================================================================================

struct dugks_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int D = 2, nc = 0, nfi = 0, nf = 0, nb = 0, nw = 0, nv = 0, nvp = 0, nxi = 0, nax = 0, K = 0;
    int storage = DUGKS_STORAGE_FP64, n_cases = 1;
    bool has_h = true;
    int nt = 64, niter = 1, nib = 1, nvb = 1, vb = 64;  // k_face velocity blocks: nvb x (vb = nt * nib)
    int ntc = 64, nibc = 1, nvbc = 1;                  // k_cell velocity blocks
    int nb_cap = 1, nbc_cap = 1;                        // partial-sum buffers sized for these
    // online launch-shape tuning: candidate (nt, nib) per kernel, all tiling the same nvp
    std::vector<std::pair<int, int>> face_cand, cell_cand;
    int tuning = 0, tune_i = -1;
    std::vector<double> tune_tf, tune_tc;   // per candidate: the shortest time measured
    static constexpr int TUNE_ROUNDS = 2;   // every candidate is timed on this many steps
    int ns_max = 0;      // largest face count of a cell
    double R = 0, omega = 0, Tref = 0, mu_ref = 0, Pr = 0, xi_max = 0;
    std::vector<int> perm;          // device velocity -> reference velocity, -1 for padding
    double off_rho = 0.0, off_T = 0.0;  // FP32 storage: rest Maxwellian of the row offsets
    double *moff = nullptr;
    std::vector<double> dcoef_int;  // internal deltaCoeffs (Courant number)
    std::vector<double> face_w;     // surfaceInterpolation weights
    std::vector<double> dt_host;    // staging of the per-case dt
    std::vector<int> patch_type, patch_start, patch_size, patch_wall0;
    std::vector<int> wall_face_slot;  // per wall: its slot index
    std::vector<int> ncell_slots;

    // device buffers
    uint32_t *vidx = nullptr, *mir = nullptr;
    int has_sym = 0;     // symmetry, far-field or fixed patches present (they need face-value mode)
    int nopen = 0;       // number of such faces
    int npress = 0;      // pressure-inlet / outlet faces among them
    double *bpress = nullptr;
    int *open = nullptr; // their face indices
    double *axw = nullptr;
    double *axis = nullptr;
    int *cs = nullptr, *ccase = nullptr, *flag = nullptr, *wfslot = nullptr;
    Slot *slots = nullptr;
    double *cV = nullptr, *dtc = nullptr, *dtp = nullptr, *muc = nullptr;
    double *crho = nullptr, *cU = nullptr, *cT = nullptr, *cq = nullptr, *ctau = nullptr;
    double *frho = nullptr, *fU = nullptr, *fT = nullptr, *fq = nullptr;
    FaceEq *feq = nullptr;
    WallC *wall = nullptr;
    double *wexp = nullptr, *qwall = nullptr, *swall = nullptr;
    CellSt *cst = nullptr;
    FaceX *fx = nullptr;
    int *corder = nullptr, *forder = nullptr;
    int *mxf = nullptr, *wallf = nullptr;   // k_face_moments_mx tiles; wall faces
    int nmx_tiles = 0, nwallf = 0;
    int face_mfma = 1;   // FP64 face-value mode: face moments on the matrix cores
    double *fpart = nullptr, *cpart = nullptr;
    void *gb[2] = {nullptr, nullptr}, *hb[2] = {nullptr, nullptr};
    void *gB[2] = {nullptr, nullptr}, *hB[2] = {nullptr, nullptr};
    int face_values = 0;  // face-value mode: the face values are stored, k_cell reads them
    int face_pass_legacy = 0;  // face-value mode with the single face pass (k_face storing)
    int fuse_s1 = 1;           // one workgroup per cell: k_cell also forms the macros and S1
    void *fvg = nullptr, *fvh = nullptr;
    int cur = 0;         // buffer holding the current state
    int raw = 1;         // state holds g~ (1) or gbar prepared with dt_prep (0)
    double *df_scratch = nullptr;            // this device's copy of `scratch` (dugks_get_df_noalloc)
    hipError_t df_scratch_status = hipSuccess;
    int initialised = 0;
    std::vector<double> dt_prep;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // captured two-step graphs, one per starting ping-pong side (kernel arguments hold the
    // buffer pointers of that side; dt is read from device memory, so no re-capture on change)
    int use_graphs = 1;
    hipGraphExec_t g2[2] = {nullptr, nullptr};
    hipEvent_t tev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double last_ms = 0.0;
    int profiling = 0;
    std::vector<hipEvent_t> pev;     // NKERN + 1 events per profiled step
    double prof[8] = {0, 0, 0, 0, 0, 0, 0, 0};
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
    p.nib = h->nib; p.nb = h->nvb; p.vb = h->vb;
    p.ntc = h->ntc; p.nibc = h->nibc; p.nbc = h->nvbc; p.vbc = h->ntc * h->nibc;
    p.R = h->R; p.Pr = h->Pr; p.omega = h->omega; p.Tref = h->Tref; p.kc = (double)(h->K + 3 - h->D);
    p.vidx = h->vidx; p.mir = h->mir; p.axw = h->axw; p.moff = h->moff;
    p.axis = h->axis; p.cs = h->cs; p.slots = h->slots; p.cV = h->cV; p.ccase = h->ccase;
    p.dtc = h->dtc; p.muc = h->muc;
    p.crho = h->crho; p.cU = h->cU; p.cT = h->cT; p.cq = h->cq; p.ctau = h->ctau;
    p.frho = h->frho; p.fU = h->fU; p.fT = h->fT; p.fq = h->fq; p.feq = h->feq;
    p.wall = h->wall; p.wexp = h->wexp; p.qwall = h->qwall; p.swall = h->swall;
    p.cst = h->cst; p.fx = h->fx;
    p.corder = h->corder; p.forder = h->forder; p.fpart = h->fpart; p.cpart = h->cpart;
    p.gb_in = (const TS *)h->gb[h->cur]; p.hb_in = (const TS *)h->hb[h->cur];
    p.gb_out = (TS *)h->gb[1 - h->cur]; p.hb_out = (TS *)h->hb[1 - h->cur];
    p.gBg_in = (const TS *)h->gB[h->cur]; p.gBh_in = (const TS *)h->hB[h->cur];
    p.gBg_out = (TS *)h->gB[1 - h->cur]; p.gBh_out = (TS *)h->hB[1 - h->cur];
    p.nopen = h->nopen;
    p.bpress = h->bpress;
    p.open = h->open;
    p.fsel = nullptr;
    p.nfsel = 0;
    p.mxf = h->mxf;
    p.fvg = h->face_values ? (TS *)h->fvg : nullptr;
    p.fvh = h->face_values ? (TS *)h->fvh : nullptr;
    p.flag = h->flag;
    return p;
}

// ---- launch dispatch over (D, H, TS, NIT) ----


// dynamic LDS of k_cell with stencil template NS: face tables (2 NS nax), velocity table
// (3 nax + 2), tabC (nax)
size_t cell_lds(const dugks_engine *h, int ns) { return sizeof(double) * ((4 + 2 * (size_t)ns) * h->nax + 2); }

dim3 block_grid(int n, int nb) { return dim3((unsigned)(((n + NXCD - 1) / NXCD) * NXCD), (unsigned)nb); }

// Kernel slots of a step (profiling and tuning events): face pass 1 (k_face, or k_face_values in
// face-value mode), face pass 2 (k_face_moments in face-value mode; empty otherwise), k_face_fin,
// k_cell, k_cell_fin, k_cell_s1.
constexpr int NKERN = 6;
constexpr int S1_NT = 256;  // workgroup size of k_cell_s1

template <int D, bool H, int NS, typename TS, bool FV>
void launch_step_fv(dugks_engine *h, hipEvent_t *ev)
{
    P<TS> p = params<TS>(h);
    if (ev) (void)hipEventRecord(ev[0], h->stream);
    if (FV && !h->face_pass_legacy) {
        hipLaunchKernelGGL((k_face_values<D, H, NS, TS>), block_grid(h->nc, h->nvb), dim3(h->nt), 0, h->stream, p);
        if (h->nopen)
            hipLaunchKernelGGL((k_face_open<D, H, NS, TS>), dim3(h->nopen, h->nvb), dim3(h->nt), 0, h->stream, p);
        if (ev) (void)hipEventRecord(ev[1], h->stream);
        bool done = false;
        {
            if (h->face_mfma) {
                // non-wall faces on the matrix cores, wall faces (outgoing sums) on the vector units
                if (h->nmx_tiles)
                    hipLaunchKernelGGL((k_face_moments_mx<D, H, TS>), dim3(h->nmx_tiles), dim3(MX_NT), 0, h->stream, p);
                if (h->nwallf) {
                    P<TS> pw = p;
                    pw.fsel = h->wallf;
                    pw.nfsel = h->nwallf;
                    hipLaunchKernelGGL((k_face_moments<D, H, TS>), block_grid(h->nwallf, h->nvb), dim3(h->nt), 0, h->stream, pw);
                }
                done = true;
            }
        }
        if (!done)
            hipLaunchKernelGGL((k_face_moments<D, H, TS>), block_grid(h->nf, h->nvb), dim3(h->nt), 0, h->stream, p);
    } else {
        hipLaunchKernelGGL((k_face<D, H, NS, TS, FV>), block_grid(h->nf, h->nvb), dim3(h->nt), 0, h->stream, p);
        if (ev) (void)hipEventRecord(ev[1], h->stream);
    }
    if (ev) (void)hipEventRecord(ev[2], h->stream);
    hipLaunchKernelGGL((k_face_fin<D, TS>), dim3((h->nf + 255) / 256), dim3(256), 0, h->stream, p);
    if (ev) (void)hipEventRecord(ev[3], h->stream);
    if constexpr (FV) {
        if (h->fuse_s1 && h->nvbc == 1) {
            // k_cell with the cell macros and the next step's S1 in its tail (k_cell_fin and
            // k_cell_s1 fused; their profile slots stay empty)
            hipLaunchKernelGGL((k_cell<D, H, NS, TS, FV, true>), block_grid(h->nc, 1), dim3(h->ntc), cell_lds(h, NS),
                               h->stream, p);
            if (ev) (void)hipEventRecord(ev[4], h->stream);
            if (h->npress)
                hipLaunchKernelGGL((k_pressure_bc<D, TS>), dim3((h->nopen + 255) / 256), dim3(256), 0, h->stream, p);
            if (ev) (void)hipEventRecord(ev[5], h->stream);
            if (ev) (void)hipEventRecord(ev[6], h->stream);
            return;
        }
    }
    hipLaunchKernelGGL((k_cell<D, H, NS, TS, FV>), block_grid(h->nc, h->nvbc), dim3(h->ntc), cell_lds(h, NS), h->stream, p);
    if (ev) (void)hipEventRecord(ev[4], h->stream);
    hipLaunchKernelGGL((k_cell_fin<D, TS>), dim3((h->nc + 255) / 256), dim3(256), 0, h->stream, p);
    if (h->npress)
        hipLaunchKernelGGL((k_pressure_bc<D, TS>), dim3((h->nopen + 255) / 256), dim3(256), 0, h->stream, p);
    if (ev) (void)hipEventRecord(ev[5], h->stream);
    hipLaunchKernelGGL((k_cell_s1<D, H, TS>), dim3(h->nc), dim3(S1_NT), 0, h->stream, p);
    if (ev) (void)hipEventRecord(ev[6], h->stream);
}

template <int D, bool H, int NS, typename TS>
void launch_step_ns(dugks_engine *h, hipEvent_t *ev)
{
    if (h->face_values) launch_step_fv<D, H, NS, TS, true>(h, ev);
    else launch_step_fv<D, H, NS, TS, false>(h, ev);
}

// stencil sizes: 4 in 1-D (2 faces per cell), 4 (triangles, quads) or 8 in 2-D, 6 (hexahedra,
// tetrahedra) or 8 in 3-D, and 16 for general polygons and polyhedra
template <int D, bool H, typename TS>
void launch_step_t(dugks_engine *h, hipEvent_t *ev)
{
    if constexpr (D == 1) {
        launch_step_ns<D, H, 4, TS>(h, ev);
    } else {
        if (h->ns_max <= (D == 2 ? 4 : 6)) launch_step_ns<D, H, D == 2 ? 4 : 6, TS>(h, ev);
        else if (h->ns_max <= 8) launch_step_ns<D, H, 8, TS>(h, ev);
        else launch_step_ns<D, H, MAXS, TS>(h, ev);
    }
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
    } else if (h->D == 2) {
        // in 2-D, K + 3 - D >= 1: h is always carried
        if (fp32) f.template operator()<2, true, float>(); else f.template operator()<2, true, double>();
    } else {
        // 1-D (fvDVM.C:198-215): K + 3 - D >= 2, h carried
        if (fp32) f.template operator()<1, true, float>(); else f.template operator()<1, true, double>();
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
        // set initial rho in pressureIn/Out BC: the constructor runs S10 once on the initial
        // state before the first step (fvDVM.C:1071-1072)
        if (h->npress)
            hipLaunchKernelGGL((k_pressure_bc<D, TS>), dim3((h->nopen + 255) / 256), dim3(256), 0, h->stream, p);
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
    int ref_nv;   // > 0: rows in reference velocity order (see k_extract)
    template <int D, bool H, typename TS>
    void operator()()
    {
        P<TS> p = params<TS>(h);
        hipLaunchKernelGGL((k_extract<D, H, TS>), dim3(n), dim3(h->nt), 0, h->stream, p, cells,
                           (const double *)h->dtp, h->raw, g, hh, ref_nv);
    }
};

// Velocity blocks: nb blocks of vb = nt * nib velocities, nt a multiple of 64 up to MAXNT and
// nib <= NIBMAX.  Least padding first, then a block size close to 4096 velocities (long enough to
// amortise a workgroup's set-up and reduction, measured on the 48^3 x 24^3 cavity), then the
// workgroup size closest to 256.
// DUGKS_NT / DUGKS_NIB in the environment pin the workgroup size / iterations (tuning).
void choose_blocks(int nv, int &nt, int &nib, int &nb, int &nvp)
{
    long best_pad = 1L << 40;
    double best_sz = 1e30, best_wg = 1e30;
    const char *ent = std::getenv("DUGKS_NT"), *eit = std::getenv("DUGKS_NIB");
    const int pin_t = ent ? std::atoi(ent) : 0, pin_it = eit ? std::atoi(eit) : 0;
    for (int t = 64; t <= MAXNT; t += 64)
        for (int it = 1; it <= NIBMAX; ++it) {
            if ((pin_t > 0 && t != pin_t) || (pin_it > 0 && it != pin_it)) continue;
            const int vb = t * it;
            const int b = (nv + vb - 1) / vb;
            const long pad = (long)b * vb - nv;
            const double sz = std::fabs(std::log((double)vb / 4096.0));
            const double wg = std::fabs(t - 256.0);
            const bool better = pad < best_pad || (pad == best_pad && (sz < best_sz - 1e-12 ||
                                                  (std::fabs(sz - best_sz) <= 1e-12 && wg < best_wg)));
            if (better) {
                best_pad = pad;
                best_sz = sz;
                best_wg = wg;
                nt = t;
                nib = it;
                nb = b;
            }
        }
    nvp = nb * nt * nib;
}

// Launch order: Morton (Z-order) key of the centres, case-major for a batch of cases.
std::vector<int> morton_order(int n, const double *X, const int *group)
{
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], X[3 * i + d]);
            hi[d] = std::max(hi[d], X[3 * i + d]);
        }
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; ++i) {
        uint64_t q[3];
        for (int d = 0; d < 3; ++d) {
            const double ext = hi[d] - lo[d];
            const double u = ext > 0 ? (X[3 * i + d] - lo[d]) / ext : 0.0;
            q[d] = (uint64_t)std::min(2097151.0, std::max(0.0, std::floor(u * 2097151.0)));
        }
        key[i] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        const int ga = group ? group[a] : 0, gb = group ? group[b] : 0;
        if (ga != gb) return ga < gb;
        return key[a] < key[b];
    });
    return ord;
}

// Launch shapes that tile the padded velocity count: (nt, nib) with nt a multiple of 64 up to
// MAXNT, nib <= NIBMAX, nt * nib dividing nvp into at most 16 blocks.  The cell kernel's
// default is one workgroup per cell (the largest nt that tiles nvp in NIBMAX iterations or
// fewer), measured faster than the face kernel's blocking on the 48^3 x 24^3 cavity; the face
// kernel keeps choose_blocks' shape.  Partial-sum buffers are sized for the largest block count.
void set_face_shape(dugks_engine *h, int nt, int nib)
{
    h->nt = nt;
    h->nib = nib;
    h->vb = nt * nib;
    h->nvb = h->nvp / h->vb;
    h->niter = h->nvp / nt;
}

void set_cell_shape(dugks_engine *h, int nt, int nib)
{
    h->ntc = nt;
    h->nibc = nib;
    h->nvbc = h->nvp / (nt * nib);
}

// Default cell-kernel shape: one workgroup per cell, the largest workgroup that tiles nvp in
// NIBMAX iterations or fewer, up to 256 threads in face-value mode and MAXNT otherwise (the
// fastest on the 48^3 x 24^3 cavity in each mode); else the face kernel's shape.
void default_cell_shape(dugks_engine *h)
{
    int ct = 0;
    for (int t = h->face_values ? 256 : MAXNT; t >= 64 && !ct; t -= 64)
        if (h->nvp % t == 0 && h->nvp / t <= NIBMAX) ct = t;
    for (int t = MAXNT; t >= 64 && !ct; t -= 64)
        if (h->nvp % t == 0 && h->nvp / t <= NIBMAX) ct = t;
    if (ct) set_cell_shape(h, ct, h->nvp / ct);
    else set_cell_shape(h, h->nt, h->nib);
}

void shape_candidates(dugks_engine *h)
{
    h->face_cand.clear();
    for (int nb = 1; nb <= 16; ++nb) {
        if (h->nvp % nb) continue;
        const int vb = h->nvp / nb;
        for (int t = 64; t <= MAXNT; t += 64) {
            if (vb % t || vb / t > NIBMAX) continue;
            h->face_cand.emplace_back(t, vb / t);
        }
    }
    h->face_cand.emplace_back(h->nt, h->nib);   // choose_blocks' shape, always a candidate
    std::sort(h->face_cand.begin(), h->face_cand.end());
    h->face_cand.erase(std::unique(h->face_cand.begin(), h->face_cand.end()), h->face_cand.end());
    h->cell_cand = h->face_cand;
    default_cell_shape(h);
    h->nb_cap = h->nvb;
    h->nbc_cap = h->nvbc;
    for (auto &c : h->face_cand) h->nb_cap = std::max(h->nb_cap, h->nvp / (c.first * c.second));
    h->nbc_cap = h->nb_cap = std::max(h->nb_cap, h->nbc_cap);
}

using dugks_lsq::mag3;

// engine face kind of a non-cyclic boundary patch type
int kind_of(int patch_type)
{
    switch (patch_type) {
    case DUGKS_PATCH_WALL: return KIND_WALL;
    case DUGKS_PATCH_SYMMETRY: return KIND_SYM;
    case DUGKS_PATCH_SYMMETRY_PLANE: return KIND_SPL;
    case DUGKS_PATCH_ZEROGRADIENT: return KIND_ZG;
    default: return KIND_FAR;   // far field, fixed, pressure inlet / outlet
    }
}

int build(dugks_engine *h, const dugks_case *cs)
{
    const int nc = cs->n_cells, nfi = cs->n_internal_faces, nf = cs->n_faces, nb = nf - nfi;
    h->D = cs->dim; h->nc = nc; h->nfi = nfi; h->nf = nf; h->nb = nb;
    h->nxi = cs->n_xi; h->K = cs->K; h->R = cs->R; h->omega = cs->omega; h->Tref = cs->T_ref;
    h->mu_ref = cs->mu_ref; h->Pr = cs->Pr; h->xi_max = cs->xi_max; h->storage = cs->storage;
    h->n_cases = cs->n_cases > 1 ? cs->n_cases : 1;
    h->has_h = (cs->K + 3 - cs->dim) != 0;
    h->nv = cs->dim == 3 ? cs->n_xi * cs->n_xi * cs->n_xi : (cs->dim == 2 ? cs->n_xi * cs->n_xi : cs->n_xi);
    h->nax = cs->dim * cs->n_xi;
    const int D = h->D;

    // ---- boundary bookkeeping ----
    std::vector<int> bkind(nb, 0), bpartner(nb, -1), bpatch(nb, -1), bwall(nb, -1), bsym(nb, -1);
    std::vector<double> Ub(3 * (size_t)nb, 0.0), Tb(nb, 0.0), rhob(nb, 0.0), pressb(nb, 0.0);
    h->patch_type.assign(cs->n_patches, 0);
    h->patch_start.assign(cs->n_patches, 0);
    h->patch_size.assign(cs->n_patches, 0);
    h->patch_wall0.assign(cs->n_patches, -1);
    int nw = 0;
    for (int pi = 0; pi < cs->n_patches; ++pi) {
        const dugks_patch &pp = cs->patches[pi];
        if (pp.start < nfi || pp.size < 0 || pp.start + pp.size > nf)
            return fail(h, DUGKS_ERR_INVALID, "patch " + std::to_string(pi) + " face range outside the boundary");
        const bool symp = pp.type == DUGKS_PATCH_SYMMETRY || pp.type == DUGKS_PATCH_SYMMETRY_PLANE;
        if (pp.type != DUGKS_PATCH_WALL && pp.type != DUGKS_PATCH_CYCLIC && !symp &&
            pp.type != DUGKS_PATCH_FARFIELD && pp.type != DUGKS_PATCH_FIXED &&
            pp.type != DUGKS_PATCH_PRESSURE_IN && pp.type != DUGKS_PATCH_PRESSURE_OUT &&
            pp.type != DUGKS_PATCH_ZEROGRADIENT)
            return fail(h, DUGKS_ERR_UNSUPPORTED, "patch " + std::to_string(pi) + ": unsupported type");
        const bool press = pp.type == DUGKS_PATCH_PRESSURE_IN || pp.type == DUGKS_PATCH_PRESSURE_OUT;
        const bool far = pp.type == DUGKS_PATCH_FARFIELD || pp.type == DUGKS_PATCH_FIXED || press;
        if (far) {
            if (!pp.T || (!press && !pp.rho) || (pp.type == DUGKS_PATCH_FIXED && !pp.U))
                return fail(h, DUGKS_ERR_INVALID, "open patch without rho, T (and U)");
            if (press && !(pp.pressure > 0)) return fail(h, DUGKS_ERR_INVALID, "pressure patch pressure must be > 0");
            h->has_sym = 1;
        }
        h->patch_type[pi] = pp.type; h->patch_start[pi] = pp.start; h->patch_size[pi] = pp.size;
        int sym_axis = -1;
        if (pp.type == DUGKS_PATCH_WALL) {
            if (!pp.U || !pp.T) return fail(h, DUGKS_ERR_INVALID, "wall patch without U/T");
            h->patch_wall0[pi] = nw;
        } else if (symp) {
            // the mirror axis comes from the patch's first face normal (discreteVelocity.C:766-768,
            // 801-803)
            h->has_sym = 1;
            if (pp.size) {
                const double *S0 = cs->Sf + 3 * (size_t)pp.start;
                const double m = mag3(S0);
                sym_axis = 0;
                for (int d = 1; d < 3; ++d)
                    if (std::fabs(S0[d]) > std::fabs(S0[sym_axis])) sym_axis = d;
                if (!(m > 0) || !(std::fabs(S0[sym_axis]) / m > 1.0 - 1e-9) || sym_axis >= cs->dim)
                    return fail(h, DUGKS_ERR_UNSUPPORTED,
                                "symmetry patch " + std::to_string(pi) + " must be normal to a solution axis");
            }
        } else if (pp.type == DUGKS_PATCH_ZEROGRADIENT) {
            h->has_sym = 1;   // its incoming half is a stored face value: face-value mode only
        } else if (pp.type == DUGKS_PATCH_CYCLIC) {
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
            } else if (symp) {
                // symmetryMod: a fixedGradient row like a wall's; symmetryPlane: a record whose row
                // is never used (phi_b = phi_P); the wall constants are placeholders
                Tb[b] = cs->T_ref;
                bsym[b] = sym_axis;
                bwall[b] = nw++;
            } else if (pp.type == DUGKS_PATCH_ZEROGRADIENT) {
                // a fixedGradient row (discreteVelocity.C:436-470); wall constants are placeholders
                Tb[b] = cs->T_ref;
                bwall[b] = nw++;
            } else if (far) {
                // a fixedGradient row; the wall record carries rho_b, T_b (and U_b)
                if (pp.type == DUGKS_PATCH_FIXED)
                    for (int d = 0; d < 3; ++d) Ub[3 * b + d] = pp.U[3 * i + d];
                Tb[b] = pp.T[i];
                // pressureIn / pressureOut rho patch values start at one (pIn.C:75); S10 in
                // dugks_init_equilibrium replaces them before the first step
                rhob[b] = press ? 1.0 : pp.rho[i];
                pressb[b] = press ? pp.pressure : 0.0;
                if (!(Tb[b] > 0) || !(rhob[b] > 0)) return fail(h, DUGKS_ERR_INVALID, "boundary rho and T must be > 0");
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
    std::vector<double> pvec, nvec, magSf(nf);
    for (int f = 0; f < nf; ++f) magSf[f] = mag3(cs->Sf + 3 * f);
    if (const char *why = dugks_lsq::vectors(cs, bkind, pvec, nvec)) return fail(h, DUGKS_ERR_INVALID, why);
    const double *Cc = cs->C, *w = cs->weights;

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
            return fail(h, DUGKS_ERR_UNSUPPORTED, "cell " + std::to_string(c) + " has more than 16 faces");
        if (D == 1 && ns > 2)
            return fail(h, DUGKS_ERR_INVALID, "1-D mesh: cell " + std::to_string(c) + " has more than 2 faces");
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
                if (bkind[b] != DUGKS_PATCH_CYCLIC) {
                    s.kind = kind_of(bkind[b]);
                    s.bidx = bwall[b];
                    if (s.kind == KIND_SYM || s.kind == KIND_SPL) s.other = bsym[b];
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
    // ---- per-face data of k_face ----
    std::vector<FaceX> fxs(nf);
    std::vector<int> open_faces;
    for (int f = 0; f < nf; ++f) {
        FaceX x;
        std::memset(&x, 0, sizeof x);
        const int o = cs->owner[f];
        for (int d = 0; d < 3; ++d) {
            x.S[d] = cs->Sf[3 * f + d];
            x.dP[d] = cs->Cf[3 * f + d] - Cc[3 * o + d];
        }
        x.P = o;
        x.N = -1;
        x.kind = KIND_OWN;
        x.bidx = -1;
        if (f < nfi) {
            x.N = cs->neighbour[f];
            for (int d = 0; d < 3; ++d) x.dN[d] = cs->Cf[3 * f + d] - Cc[3 * x.N + d];
        } else {
            const int b = f - nfi;
            if (bkind[b] != DUGKS_PATCH_CYCLIC) {
                x.kind = kind_of(bkind[b]);
                x.bidx = bwall[b];
                if (x.kind == KIND_SYM || x.kind == KIND_SPL) x.N = bsym[b];
                // incoming halves written by k_face_open
                if (x.kind == KIND_SYM || x.kind == KIND_SPL || x.kind == KIND_FAR) open_faces.push_back(f);
                for (int d = 0; d < 3; ++d) x.dN[d] = cs->Sf[3 * f + d] / magSf[f];
            } else {
                x.kind = KIND_CYC;
                x.bidx = b;
                x.N = cs->owner[nfi + bpartner[b]];
                x.w = w[f];
            }
        }
        fxs[f] = x;
    }

    // ---- velocities: tensor product in reference order, then grouped by sign pattern ----
    const int nxi = h->nxi;
    if (nxi < 2 || nxi > 64) return fail(h, DUGKS_ERR_UNSUPPORTED, "n_xi must be in [2, 64]");
    if (h->nax > NAXMAX) return fail(h, DUGKS_ERR_UNSUPPORTED, "velocity table too large");
    const int nv = h->nv;
    std::vector<double> rx(nv), ry(nv), rz(nv);
    std::vector<int> rid(nv);
    {
        int i = 0;
        const double *X = cs->xi;
        if (D == 3) {
            for (int iz = 0; iz < nxi; ++iz)
                for (int iy = 0; iy < nxi; ++iy)
                    for (int ix = 0; ix < nxi; ++ix) {
                        rx[i] = X[ix]; ry[i] = X[iy]; rz[i] = X[iz];
                        rid[i] = ix | (iy << 8) | (iz << 16);
                        ++i;
                    }
        } else if (D == 2) {
            for (int iy = 0; iy < nxi; ++iy)
                for (int ix = 0; ix < nxi; ++ix) {
                    rx[i] = X[ix]; ry[i] = X[iy]; rz[i] = 0.0;
                    rid[i] = ix | (iy << 8);
                    ++i;
                }
        } else {
            for (int ix = 0; ix < nxi; ++ix) {
                rx[i] = X[ix]; ry[i] = 0.0; rz[i] = 0.0;
                rid[i] = ix;
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
    choose_blocks(nv, h->nt, h->nib, h->nvb, h->nvp);
    h->vb = h->nt * h->nib;
    h->niter = h->nvb * h->nib;
    shape_candidates(h);
    if (h->nt <= 0 || h->nt > MAXNT || h->nvp < nv)
        return fail(h, DUGKS_ERR_UNSUPPORTED, "velocity blocking failed");
    // one dispatch holds at most 2^32 - 1 work-items per grid dimension
    if ((double)((std::max(nc, nf) + NXCD - 1) / NXCD * NXCD) * MAXNT >= 4294967295.0)
        return fail(h, DUGKS_ERR_UNSUPPORTED, "mesh too large for one dispatch per kernel");
    // ---- least-squares stencils (entry j = slot j; rows: cells, then walls) ----
    std::vector<CellSt> cst(nc);
    h->ns_max = 0;
    for (int c = 0; c < nc; ++c) {
        CellSt t;
        std::memset(&t, 0, sizeof t);
        const int ns = cnt[c + 1] - cnt[c];
        h->ns_max = std::max(h->ns_max, ns);
        t.inner = ns > 0;
        for (int j = 0; j < ns; ++j) {
            const int kd = sl[cnt[c] + j].kind;
            if (kd != KIND_OWN && kd != KIND_NEI) t.inner = 0;
        }
        for (int j = 0; j < MAXS; ++j) {
            t.off[j] = (int64_t)c * h->nvp;
            if (j >= ns) continue;
            const Slot &q = sl[cnt[c] + j];
            for (int d = 0; d < 3; ++d) t.L[j][d] = q.L[d];
            if (grad_row(q.kind)) {
                t.off[j] = (int64_t)(nc + q.bidx) * h->nvp;
                t.a[j] = q.aux;
                t.wmask |= 1 << j;
            } else if (q.kind == KIND_SPL) {
                // symmetryPlane: phi_b = phi_P, so the entry reads the cell's own row and adds
                // L (phi_P - phi_P) = 0 (leastSquaresGrad's boundary term)
                t.off[j] = (int64_t)c * h->nvp;
            } else {
                t.off[j] = (int64_t)q.other * h->nvp;
            }
        }
        cst[c] = t;
    }
    const int nvp = h->nvp;
    std::vector<uint32_t> vidx(nvp, VPAD);
    h->perm.assign(nvp, -1);
    for (int k = 0; k < nv; ++k) {
        const int r = order[k];
        vidx[k] = (uint32_t)rid[r];
        h->perm[k] = r;
    }
    // mirrored velocities (symmXtgID / symmYtgID / symmZtgID, fvDVM.C:152-166, 183-195) in device order
    std::vector<uint32_t> mir(3 * (size_t)nvp);
    {
        std::vector<int> inv(nv);
        for (int k = 0; k < nv; ++k) inv[h->perm[k]] = k;
        for (int k = 0; k < nvp; ++k)
            for (int a = 0; a < 3; ++a) mir[(size_t)a * nvp + k] = (uint32_t)k;
        for (int k = 0; k < nv; ++k) {
            const int r = h->perm[k];
            const int ix = r % nxi, iy = D >= 2 ? (r / nxi) % nxi : 0, iz = D == 3 ? r / (nxi * nxi) : 0;
            const int ny = D >= 2 ? nxi : 1, nz = D == 3 ? nxi : 1;
            mir[k] = (uint32_t)inv[(iz * ny + iy) * nxi + (nxi - 1 - ix)];
            mir[(size_t)nvp + k] = (uint32_t)inv[(iz * ny + (ny - 1 - iy)) * nxi + ix];
            mir[2 * (size_t)nvp + k] = (uint32_t)inv[((nz - 1 - iz) * ny + iy) * nxi + ix];
        }
    }
    std::vector<double> axis(h->nax), axw(h->nax);
    for (int a = 0; a < D; ++a)
        for (int i = 0; i < nxi; ++i) {
            axis[a * nxi + i] = cs->xi[i];
            axw[a * nxi + i] = cs->xi_weights[i];
        }

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
    // ---- launch orders ----
    const std::vector<int> corder = morton_order(nc, Cc, ccase.data());
    std::vector<int> fcase(nf);
    for (int f = 0; f < nf; ++f) fcase[f] = ccase[cs->owner[f]];
    const std::vector<int> forder = morton_order(nf, cs->Cf, fcase.data());
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
    std::vector<double> wpress(nw, 0.0);
    for (int b = 0; b < nb; ++b) {
        if (bwall[b] < 0) continue;
        WallC W;
        std::memset(&W, 0, sizeof W);
        for (int d = 0; d < 3; ++d) W.U[d] = Ub[3 * b + d];
        W.T = Tb[b];
        double tp = 2.0 * PI * (cs->R * Tb[b]);
        W.pre = D == 3 ? 1.0 / (tp * std::sqrt(tp)) : (D == 2 ? 1.0 / tp : 1.0 / std::sqrt(tp));
        W.rho_b = rhob[b];
        W.mode = bkind[b] == DUGKS_PATCH_FIXED ? 1.0
               : bkind[b] == DUGKS_PATCH_PRESSURE_IN ? 2.0 : bkind[b] == DUGKS_PATCH_PRESSURE_OUT ? 3.0 : 0.0;
        wpress[bwall[b]] = pressb[b];
        if (bkind[b] == DUGKS_PATCH_PRESSURE_IN || bkind[b] == DUGKS_PATCH_PRESSURE_OUT) ++h->npress;
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
    if ((r = dalloc(h, &h->vidx, nvp)) || (r = dalloc(h, &h->mir, 3 * (size_t)nvp)) || (r = dalloc(h, &h->axw, h->nax)) || (r = dalloc(h, &h->moff, (size_t)h->nax + 2)) ||
        (r = dalloc(h, &h->axis, h->nax)) || (r = dalloc(h, &h->cs, nc + 1)) || (r = dalloc(h, &h->slots, nslot)) ||
        (r = dalloc(h, &h->cV, nc)) || (r = dalloc(h, &h->ccase, nc)) || (r = dalloc(h, &h->dtc, h->n_cases)) ||
        (r = dalloc(h, &h->dtp, h->n_cases)) || (r = dalloc(h, &h->muc, h->n_cases)) ||
        (r = dalloc(h, &h->crho, nc)) || (r = dalloc(h, &h->cU, 3 * (size_t)nc)) || (r = dalloc(h, &h->cT, nc)) ||
        (r = dalloc(h, &h->cq, 3 * (size_t)nc)) || (r = dalloc(h, &h->ctau, nc)) || (r = dalloc(h, &h->frho, nf)) ||
        (r = dalloc(h, &h->fU, 3 * (size_t)nf)) || (r = dalloc(h, &h->fT, nf)) || (r = dalloc(h, &h->fq, 3 * (size_t)nf)) ||
        (r = dalloc(h, &h->feq, nf)) ||
        (r = dalloc(h, &h->wall, nw)) || (r = dalloc(h, &h->wexp, (size_t)nw * h->nax)) ||
        (r = dalloc(h, &h->qwall, 3 * (size_t)nw)) || (r = dalloc(h, &h->swall, 9 * (size_t)nw)) ||
        (r = dalloc(h, &h->fx, nf)) || (r = dalloc(h, &h->cst, nc)) || (r = dalloc(h, &h->corder, nc)) ||
        (r = dalloc(h, &h->forder, nf)) || (r = dalloc(h, &h->fpart, (size_t)nf * h->nb_cap * (NMom<3>::v + 1))) ||
        (r = dalloc(h, &h->cpart, (size_t)nc * h->nbc_cap * NMom<3>::v)) || (r = dalloc(h, &h->flag, 1)) || (r = dalloc(h, &h->wfslot, nw)))
        return r;
    const size_t es = ts_size(h);
    // one allocation per ping-pong side: nc cell rows, then nw wall rows (boundary gradient)
    const size_t row = (size_t)nc * nvp * es, brow = (size_t)(nw ? nw : 1) * nvp * es;
    for (int i = 0; i < 2; ++i) {
        if ((r = dallocb(h, &h->gb[i], row + brow))) return r;
        h->gB[i] = (char *)h->gb[i] + row;
        if (h->has_h) {
            if ((r = dallocb(h, &h->hb[i], row + brow))) return r;
            h->hB[i] = (char *)h->hb[i] + row;
        }
    }
    if ((r = upload(h, h->vidx, vidx.data(), nvp)) || (r = upload(h, h->mir, mir.data(), 3 * (size_t)nvp)) || (r = upload(h, h->axw, axw.data(), h->nax)) ||
        (r = upload(h, h->axis, axis.data(), h->nax)) || (r = upload(h, h->cs, cnt.data(), nc + 1)) ||
        (r = upload(h, h->slots, sl.data(), nslot)) || (r = upload(h, h->cV, cs->V, nc)) ||
        (r = upload(h, h->ccase, ccase.data(), nc)) || (r = upload(h, h->muc, muc.data(), h->n_cases)) ||
        (r = upload(h, h->fx, fxs.data(), nf)) || (r = upload(h, h->cst, cst.data(), nc)) ||
        (r = upload(h, h->corder, corder.data(), nc)) || (r = upload(h, h->forder, forder.data(), nf)) || (r = upload(h, h->wfslot, wfslot.data(), nw)))
        return r;
    if (nw && (r = upload(h, h->wall, wc.data(), nw))) return r;
    if (nw && ((r = dalloc(h, &h->bpress, nw)) || (r = upload(h, h->bpress, wpress.data(), nw)))) return r;
    h->nopen = (int)open_faces.size();
    if (h->nopen && ((r = dalloc(h, &h->open, open_faces.size())) || (r = upload(h, h->open, open_faces.data(), open_faces.size()))))
        return r;
    // face lists of the matrix-core face moments: non-wall faces in launch order, 16 per tile
    // (15 and an offset row with FP32 storage; -1 pads), and the wall faces, whose
    // outgoing-only sums stay on k_face_moments
    {
        std::vector<int> nw_faces, wl, mx;
        for (int i = 0; i < nf; ++i) (fxs[forder[i]].kind == KIND_WALL ? wl : nw_faces).push_back(forder[i]);
        const int per = h->storage == DUGKS_STORAGE_FP32 ? 15 : 16;
        h->nmx_tiles = (int)((nw_faces.size() + per - 1) / per);
        mx.assign((size_t)h->nmx_tiles * 16, -1);
        for (size_t t = 0; t < nw_faces.size(); ++t) mx[(t / per) * 16 + t % per] = nw_faces[t];
        h->nwallf = (int)wl.size();
        if (h->nmx_tiles && ((r = dalloc(h, &h->mxf, mx.size())) || (r = upload(h, h->mxf, mx.data(), mx.size()))))
            return r;
        if (h->nwallf && ((r = dalloc(h, &h->wallf, wl.size())) || (r = upload(h, h->wallf, wl.data(), wl.size()))))
            return r;
    }
    HIPCHK(h, hipMemset(h->flag, 0, sizeof(int)));
    h->dt_prep.assign(h->n_cases, 0.0);
    return 0;
}

void drop_graphs(dugks_engine *h)
{
    for (hipGraphExec_t &g : h->g2) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
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
    dfree(h->vidx); dfree(h->mir); dfree(h->open); dfree(h->bpress); dfree(h->axw); dfree(h->moff); dfree(h->axis);
    dfree(h->cs); dfree(h->ccase); dfree(h->flag); dfree(h->wfslot); dfree(h->slots); dfree(h->cV);
    dfree(h->dtc); dfree(h->dtp); dfree(h->muc); dfree(h->crho); dfree(h->cU); dfree(h->cT); dfree(h->cq);
    dfree(h->ctau); dfree(h->frho); dfree(h->fU); dfree(h->fT); dfree(h->fq); dfree(h->feq);
    dfree(h->wall); dfree(h->wexp); dfree(h->qwall); dfree(h->swall); dfree(h->fx); dfree(h->cst);
    dfree(h->corder); dfree(h->forder); dfree(h->fpart); dfree(h->cpart); dfree(h->mxf); dfree(h->wallf);
    for (int i = 0; i < 2; ++i) {
        dfree(h->gb[i]); dfree(h->hb[i]);
        h->gB[i] = h->hB[i] = nullptr;
    }
    dfree(h->fvg); dfree(h->fvh);
    h->face_values = 0;
    drop_graphs(h);
    for (hipEvent_t &e : h->tev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
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

// Row offsets of the FP32 storage: the rest Maxwellian g = rho (2 pi R T)^(-D/2) exp(-|xi|^2 / 2RT)
// and h = g R T (K + 3 - D) at (off_rho, off_T), as per-axis factors exp(-x_i^2 / 2RT) [nax] and
// the prefactors of g and h; all zero for FP64 storage.
int upload_offsets(dugks_engine *h)
{
    std::vector<double> m((size_t)h->nax + 2, 0.0);
    if (h->storage == DUGKS_STORAGE_FP32) {
        const double RT = h->R * h->off_T, tp = 2.0 * PI * RT;
        const double pre = h->off_rho / (h->D == 3 ? tp * std::sqrt(tp) : (h->D == 2 ? tp : std::sqrt(tp)));
        std::vector<double> ax((size_t)h->nax);
        HIPCHK(h, hipMemcpy(ax.data(), h->axis, sizeof(double) * h->nax, hipMemcpyDeviceToHost));
        for (int t = 0; t < h->nax; ++t) m[t] = std::exp(-(ax[t] * ax[t]) / (2.0 * RT));
        m[h->nax] = pre;
        m[h->nax + 1] = h->has_h ? pre * RT * (h->K + 3 - h->D) : 0.0;
    }
    return upload(h, h->moff, m.data(), m.size());
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
        // the boundary-gradient rows travel with the cell rows
        const size_t brow = (size_t)(h->nw ? h->nw : 1) * h->nvp * ts_size(h);
        HIPCHK(h, hipMemcpyAsync(h->gB[1 - h->cur], h->gB[h->cur], brow, hipMemcpyDeviceToDevice, h->stream));
        if (h->has_h)
            HIPCHK(h, hipMemcpyAsync(h->hB[1 - h->cur], h->hB[h->cur], brow, hipMemcpyDeviceToDevice, h->stream));
        h->cur = 1 - h->cur;
        HIPCHK(h, hipGetLastError());
        h->raw = 0;
        for (int i = 0; i < h->n_cases; ++i) h->dt_prep[i] = dtc[i];
        HIPCHK(h, hipMemcpyAsync(h->dtp, h->dt_host.data(), sizeof(double) * h->n_cases, hipMemcpyHostToDevice, h->stream));
    }
    if (h->profiling) {
        while (h->pev.size() < (size_t)(NKERN + 1) * nsteps) {
            hipEvent_t e;
            HIPCHK(h, hipEventCreate(&e));
            h->pev.push_back(e);
        }
    }
    int s0 = 0;
    // online launch-shape tuning: each step runs the next candidate shape of k_face and of
    // k_cell (after one warm-up step), timed with events; every candidate is timed in
    // TUNE_ROUNDS rounds and the one with the shortest time of each kernel is kept
    while (h->tuning && s0 < nsteps) {
        const int nfc = (int)h->face_cand.size(), ncc = (int)h->cell_cand.size();
        const int ncand = std::max(nfc, ncc);
        const int i = h->tune_i < 0 ? -1 : h->tune_i % ncand;
        if (i >= 0) {
            const auto &fc = h->face_cand[std::min(i, nfc - 1)];
            const auto &cc = h->cell_cand[std::min(i, ncc - 1)];
            set_face_shape(h, fc.first, fc.second);
            set_cell_shape(h, cc.first, cc.second);
        }
        for (hipEvent_t &e : h->tev)
            if (!e) HIPCHK(h, hipEventCreate(&e));
        StepOp op{h, h->tev};
        dispatch(h, op);
        h->cur = 1 - h->cur;
        ++s0;
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventSynchronize(h->tev[NKERN]));
        float tf = 0.f, tc = 0.f;
        HIPCHK(h, hipEventElapsedTime(&tf, h->tev[0], h->tev[2]));
        // the cell side: k_cell, k_cell_fin and k_cell_s1 (one kernel when fused)
        HIPCHK(h, hipEventElapsedTime(&tc, h->tev[3], h->tev[6]));
        h->tune_tf.resize(nfc, 0.0);
        h->tune_tc.resize(ncc, 0.0);
        if (i >= 0 && i < nfc && (h->tune_tf[i] == 0.0 || tf < h->tune_tf[i])) h->tune_tf[i] = tf;
        if (i >= 0 && i < ncc && (h->tune_tc[i] == 0.0 || tc < h->tune_tc[i])) h->tune_tc[i] = tc;
        if (++h->tune_i >= dugks_engine::TUNE_ROUNDS * ncand) {
            const int bf = (int)(std::min_element(h->tune_tf.begin(), h->tune_tf.end()) - h->tune_tf.begin());
            const int bc = (int)(std::min_element(h->tune_tc.begin(), h->tune_tc.end()) - h->tune_tc.begin());
            set_face_shape(h, h->face_cand[bf].first, h->face_cand[bf].second);
            set_cell_shape(h, h->cell_cand[bc].first, h->cell_cand[bc].second);
            h->tuning = 0;
            drop_graphs(h);
        }
    }
    if (!h->profiling && h->use_graphs && nsteps - s0 >= 2) {
        // replay a captured pair of steps; an odd last step runs eagerly below
        hipGraphExec_t &ex = h->g2[h->cur];
        if (!ex) {
            const int cur0 = h->cur;
            hipGraph_t g = nullptr;
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < 2; ++i) {
                StepOp op{h, nullptr};
                dispatch(h, op);
                h->cur = 1 - h->cur;
            }
            const hipError_t ce = hipStreamEndCapture(h->stream, &g);
            h->cur = cur0;
            if (ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(h, DUGKS_ERR_DEVICE, std::string("step graph capture: ") + hipGetErrorString(ce));
            }
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                ex = nullptr;
                return fail(h, DUGKS_ERR_DEVICE, std::string("step graph instantiate: ") + hipGetErrorString(ie));
            }
        }
        for (; s0 + 2 <= nsteps; s0 += 2) HIPCHK(h, hipGraphLaunch(ex, h->stream));
    }
    for (int s = s0; s < nsteps; ++s) {
        StepOp op{h, h->profiling ? &h->pev[(NKERN + 1) * (size_t)s] : nullptr};
        dispatch(h, op);
        h->cur = 1 - h->cur;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    int r = check_flag(h);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->last_ms = ms;
    if (h->profiling) {
        for (int s = 0; s < nsteps; ++s) {
            const hipEvent_t *e = &h->pev[(NKERN + 1) * (size_t)s];
            for (int i = 0; i < NKERN; ++i) {
                float a = 0.f;
                HIPCHK(h, hipEventElapsedTime(&a, e[i], e[i + 1]));
                h->prof[i] += a;
            }
        }
        h->prof_steps += nsteps;
    }
    return r;
}

const int64_t STATE_MAGIC = 0x44554b5331LL;  // "DUKS1"

}  // namespace

extern "C" {

int dugks_abi_version(void) { return DUGKS_ABI_VERSION; }

namespace {
// Address of `scratch` on each device, resolved once per device with that device current.
std::mutex df_scratch_mu;
constexpr int DF_SCRATCH_MAX_DEVICES = 64;
double *df_scratch_dev[DF_SCRATCH_MAX_DEVICES];

hipError_t df_scratch_ptr(int device, double **out)
{
    std::lock_guard<std::mutex> lock(df_scratch_mu);
    if (device < 0 || device >= DF_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!df_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        df_scratch_dev[device] = (double *)p;
    }
    *out = df_scratch_dev[device];
    return hipSuccess;
}
}  // namespace

int dugks_create(const dugks_case *cs, dugks_engine **out)
{
    if (!out) return DUGKS_ERR_INVALID;
    *out = nullptr;
    if (!cs) return DUGKS_ERR_INVALID;
    dugks_engine *h = new (std::nothrow) dugks_engine();
    if (!h) return DUGKS_ERR_MEMORY;
    int r = 0;
    if (cs->abi_version != DUGKS_ABI_VERSION) r = fail(h, DUGKS_ERR_INVALID, "abi_version mismatch");
    else if (cs->dim < 1 || cs->dim > 3) r = fail(h, DUGKS_ERR_UNSUPPORTED, "dim must be 1, 2 or 3");
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
    if (!r) {
        // face-value mode by default when its buffers take at most half of the free device
        // memory; DUGKS_FACE_VALUES=0/1 in the environment forces it off/on
        const char *fv = std::getenv("DUGKS_FACE_VALUES");
        size_t fr = 0, tot = 0;
        const size_t need = (size_t)h->nf * h->nvp * ts_size(h) * (h->has_h ? 2 : 1);
        const bool fits = hipMemGetInfo(&fr, &tot) == hipSuccess && need <= fr / 2;
        if (h->has_sym || (fv ? std::atoi(fv) != 0 : fits)) r = dugks_set_face_values(h, 1);
        const char *fl = std::getenv("DUGKS_FACE_PASS_LEGACY");
        h->face_pass_legacy = fl && std::atoi(fl) && !h->has_sym ? 1 : 0;
        // DUGKS_FACE_MFMA=0 keeps the FP64 face moments on the vector units
        const char *fm = std::getenv("DUGKS_FACE_MFMA");
        h->face_mfma = fm && !std::atoi(fm) ? 0 : 1;
        // DUGKS_FUSE_S1=0 keeps k_cell_fin and k_cell_s1 as kernels of their own
        const char *f1 = std::getenv("DUGKS_FUSE_S1");
        h->fuse_s1 = f1 && !std::atoi(f1) ? 0 : 1;
    }
    if (!r) {
        // this device's copy of the dugks_get_df_noalloc scratch; a failure is reported by that
        // entry point only
        h->df_scratch_status = hipSetDevice(h->device);
        if (h->df_scratch_status == hipSuccess) h->df_scratch_status = df_scratch_ptr(h->device, &h->df_scratch);
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
                } else if (s.kind == KIND_SYM || s.kind == KIND_SPL || s.kind == KIND_ZG ||
                           (s.kind == KIND_FAR && wc[s.bidx].mode != 1.0)) {
                    for (int d = 0; d < 3; ++d) fU[3 * f + d] = U[3 * c + d];   // U zeroGradient there
                } else if (s.kind == KIND_FAR) {
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
    {
        std::vector<double> V(nc);
        HIPCHK(h, hipMemcpy(V.data(), h->cV, sizeof(double) * nc, hipMemcpyDeviceToHost));
        double sv = 0.0, sr = 0.0, sT = 0.0;
        for (int c = 0; c < nc; ++c) {
            sv += V[c];
            sr += V[c] * rho[c];
            sT += V[c] * T[c];
        }
        h->off_rho = sr / sv;
        h->off_T = sT / sv;
        if ((r = upload_offsets(h))) return r;
    }
    h->cur = 0;
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
        ExtractOp op{h, dcells, n, dg, dh, 0};
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

int dugks_get_df_noalloc(dugks_engine *h, int32_t n, const int32_t *cells, double *g, double *h_out)
{
    if (!h || !h->stream || !h->initialised) return h ? fail(h, DUGKS_ERR_STATE, "not initialised") : DUGKS_ERR_STATE;
    if (n < 0 || (n && !cells)) return fail(h, DUGKS_ERR_INVALID, "cells");
    for (int i = 0; i < n; ++i)
        if (cells[i] < 0 || cells[i] >= h->nc) return fail(h, DUGKS_ERR_INVALID, "cell index out of range");
    if (n == 0) return 0;
    // scratch layout: the cell list (padded to 32 doubles = 256 bytes), then g[n][nv], h[n][nv]
    const size_t cell_slots = (((size_t)n * sizeof(int32_t) + 255) / 256) * 32;
    const size_t rows = (size_t)n * h->nv;
    if (cell_slots + 2 * rows > DF_SCRATCH_DOUBLES)
        return fail(h, DUGKS_ERR_INVALID,
                    "get_df_noalloc: n * nv = " + std::to_string(rows) + " exceeds the scratch capacity (2 n nv + " +
                        std::to_string(cell_slots) + " <= " + std::to_string(DF_SCRATCH_DOUBLES) + " doubles)");
    if (h->df_scratch_status != hipSuccess || !h->df_scratch)
        return fail(h, DUGKS_ERR_DEVICE,
                    std::string("get_df_noalloc: scratch address: ") + hipGetErrorString(h->df_scratch_status));
    HIPCHK(h, hipSetDevice(h->device));
    int *dcells = (int *)h->df_scratch;
    double *dg = h->df_scratch + cell_slots, *dh = dg + rows;
    hipError_t e = hipMemcpyAsync(dcells, cells, sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        ExtractOp op{h, dcells, n, dg, dh, h->nv};
        dispatch(h, op);
        e = hipGetLastError();
    }
    if (e == hipSuccess && g) e = hipMemcpyAsync(g, dg, sizeof(double) * rows, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && h_out)
        e = hipMemcpyAsync(h_out, dh, sizeof(double) * rows, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(h, DUGKS_ERR_DEVICE, std::string("get_df_noalloc: ") + hipGetErrorString(e));
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
    n += (int64_t)h->nw * (int64_t)sizeof(WallC);   // pressure patches evolve their boundary state
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
    std::memcpy(&hdr[10], &h->off_rho, sizeof(double));   // FP32 row offsets travel with the rows
    std::memcpy(&hdr[11], &h->off_T, sizeof(double));
    std::memcpy(p, hdr, sizeof hdr);
    p += sizeof hdr;
    const size_t es = ts_size(h);
    const size_t row = (size_t)h->nc * h->nvp * es, brow = (size_t)(h->nw ? h->nw : 1) * h->nvp * es;
    HIPCHK(h, hipMemcpy(p, h->gb[h->cur], row, hipMemcpyDeviceToHost)); p += row;
    if (h->has_h) { HIPCHK(h, hipMemcpy(p, h->hb[h->cur], row, hipMemcpyDeviceToHost)); p += row; }
    HIPCHK(h, hipMemcpy(p, h->gB[h->cur], brow, hipMemcpyDeviceToHost)); p += brow;
    if (h->has_h) { HIPCHK(h, hipMemcpy(p, h->hB[h->cur], brow, hipMemcpyDeviceToHost)); p += brow; }
    const size_t nc = h->nc;
    HIPCHK(h, hipMemcpy(p, h->crho, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->cU, 24 * nc, hipMemcpyDeviceToHost)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(p, h->cT, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->cq, 24 * nc, hipMemcpyDeviceToHost)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(p, h->ctau, 8 * nc, hipMemcpyDeviceToHost)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(p, h->fU, 24 * (size_t)h->nf, hipMemcpyDeviceToHost)); p += 24 * (size_t)h->nf;
    std::memcpy(p, h->dt_prep.data(), 8 * h->n_cases);
    p += 8 * h->n_cases;
    if (h->nw) HIPCHK(h, hipMemcpy(p, h->wall, sizeof(WallC) * (size_t)h->nw, hipMemcpyDeviceToHost));
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
    HIPCHK(h, hipMemcpy(h->gB[h->cur], p, brow, hipMemcpyHostToDevice)); p += brow;
    if (h->has_h) { HIPCHK(h, hipMemcpy(h->hB[h->cur], p, brow, hipMemcpyHostToDevice)); p += brow; }
    const size_t nc = h->nc;
    HIPCHK(h, hipMemcpy(h->crho, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->cU, p, 24 * nc, hipMemcpyHostToDevice)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(h->cT, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->cq, p, 24 * nc, hipMemcpyHostToDevice)); p += 24 * nc;
    HIPCHK(h, hipMemcpy(h->ctau, p, 8 * nc, hipMemcpyHostToDevice)); p += 8 * nc;
    HIPCHK(h, hipMemcpy(h->fU, p, 24 * (size_t)h->nf, hipMemcpyHostToDevice)); p += 24 * (size_t)h->nf;
    std::memcpy(h->dt_prep.data(), p, 8 * h->n_cases);
    p += 8 * h->n_cases;
    if (h->nw) HIPCHK(h, hipMemcpy(h->wall, p, sizeof(WallC) * (size_t)h->nw, hipMemcpyHostToDevice));
    h->raw = (int)hdr[7];
    std::memcpy(&h->off_rho, &hdr[10], sizeof(double));
    std::memcpy(&h->off_T, &hdr[11], sizeof(double));
    if (int rr = upload_offsets(h)) return rr;
    HIPCHK(h, hipMemcpy(h->dtp, h->dt_prep.data(), 8 * h->n_cases, hipMemcpyHostToDevice));
    return 0;
}

int dugks_last_step_ms(dugks_engine *h, double *ms)
{
    if (!h || !ms) return DUGKS_ERR_INVALID;
    *ms = h->last_ms;
    return 0;
}

int dugks_set_autotune(dugks_engine *h, int32_t on)
{
    if (!h) return DUGKS_ERR_INVALID;
    h->tuning = on ? 1 : 0;
    h->tune_i = -1;
    h->tune_tf.clear();
    h->tune_tc.clear();
    return 0;
}

int dugks_get_face_values(dugks_engine *h, int32_t *on)
{
    if (!h || !on) return DUGKS_ERR_INVALID;
    *on = h->face_values;
    return 0;
}

int dugks_autotune_steps_left(dugks_engine *h, int32_t *steps)
{
    if (!h || !steps) return DUGKS_ERR_INVALID;
    const int n = std::max((int)h->face_cand.size(), (int)h->cell_cand.size());
    *steps = h->tuning ? dugks_engine::TUNE_ROUNDS * n - h->tune_i : 0;
    return 0;
}

int dugks_get_cell_launch(dugks_engine *h, int32_t *nt, int32_t *nib, int32_t *n_blocks)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (nt) *nt = h->ntc;
    if (nib) *nib = h->nibc;
    if (n_blocks) *n_blocks = h->nvbc;
    return 0;
}

int dugks_set_graphs(dugks_engine *h, int32_t on)
{
    if (!h) return DUGKS_ERR_INVALID;
    h->use_graphs = on ? 1 : 0;
    return 0;
}

int dugks_set_face_values(dugks_engine *h, int32_t on)
{
    if (!h || !h->stream) return h ? fail(h, DUGKS_ERR_STATE, "no device resources") : DUGKS_ERR_INVALID;
    if (!on && h->has_sym)
        return fail(h, DUGKS_ERR_UNSUPPORTED, "symmetry, far-field and fixed patches need face-value mode");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (on && !h->fvg) {
        const size_t bytes = (size_t)h->nf * h->nvp * ts_size(h);
        int r;
        if ((r = dallocb(h, &h->fvg, bytes))) return r;
        if (h->has_h && (r = dallocb(h, &h->fvh, bytes))) {
            dfree(h->fvg);
            return r;
        }
    } else if (!on) {
        dfree(h->fvg);
        dfree(h->fvh);
    }
    h->face_values = on ? 1 : 0;
    default_cell_shape(h);
    drop_graphs(h);
    return 0;
}

int dugks_get_launch(dugks_engine *h, int32_t *nt, int32_t *nib, int32_t *n_blocks, int32_t *nv_padded)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (nt) *nt = h->nt;
    if (nib) *nib = h->nib;
    if (n_blocks) *n_blocks = h->nvb;
    if (nv_padded) *nv_padded = h->nvp;
    return 0;
}

int dugks_set_profiling(dugks_engine *h, int32_t on)
{
    if (!h) return DUGKS_ERR_INVALID;
    h->profiling = on ? 1 : 0;
    for (double &x : h->prof) x = 0.0;
    h->prof_steps = 0;
    return 0;
}

int dugks_get_profile(dugks_engine *h, double *ms, int64_t *steps)
{
    if (!h) return DUGKS_ERR_INVALID;
    if (ms) {
        for (int i = 0; i < NKERN; ++i) ms[i] = h->prof[i];
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
