// adjoint_core.h -- per-point work of the discrete adjoint of the DUGKS step (every patch kind of
// the engine, FP64), written once for two executors:
//   * the HIP kernels of dugks_adjoint.hip: one workgroup per cell or face, the lanes stride over
//     the velocities, velocity sums are workgroup reductions;
//   * a host executor (plain loops) used to check the same code on the CPU.
// Every routine takes the executor X, which provides
//   X::for_k(nv, fn)   run fn(k) for the velocities this worker owns,
//   X::sum<N>(v)       replace the worker's partial sums v[N] by the point's totals,
//   X::leader()        true for the one worker that writes per-point scalars.
//
// The step is fvDVM::evolution() (fvDVM.C:1086-1108) with internal faces, Maxwell walls, cyclic
// pairs, fixed-Maxwellian, far-field, pressureIn/pressureOut, zeroGradient, symmetryMod and
// symmetryPlane patches (S1-S10 of the oracle).  State n is (g~, h~ [cell][k], the lagged fixedGradient
// rows Gg, Gh [boundary face][k], the cell macros rho, U, T, q).  Its macros are the S9 values of
// step n-1 (for n = 0 the given initial fields with q = 0).  The objective is
//   J = sum over steps n >= first of sum_f c_f (rho U)_f . S_f,  (rho U)_f = sum_k w_k xi_k gbar_f,k
// with gbar_f the face values of S4 (fvDVM.C:473-483), i.e. a weighted mass flux through faces.
// Its gradient is taken with respect to the temperature of every wall face and of every fixed
// (frozen boundary Maxwellian) face.
//
// Reverse of one step (state n -> n+1), given the adjoint of g~^{n+1}, h~^{n+1} (with the adjoint
// of the macros of n+1 already folded in) and of G^{n+1}:
//   face_adj   S8 flux -> relaxed face values; S7 relaxation (Shakhov at the face, r_f(tau_f));
//              objective seed; S6 face macros; S4/S3 wall: rho_w = out / in, incoming rho_w M(T_w)
//              -> dJ/dT_w and the outgoing face values
//   recon_adj  S2 upwind reconstruction -> gbar+ and the gradient; S2 new G = n . grad
//   grad_adj   least-squares gradient -> gbar+ and the lagged rows G^n
//   s1_adj     S1 gbar+ = (1 - r) g~ + r geq(M) -> g~^n and M^n; M^n = m(g~^n) folded into g~^n
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define ADJ_HD __host__ __device__ inline
#else
#define ADJ_HD inline
#endif

namespace adj {

constexpr double VSMALL = 1.0e-300;   // OpenFOAM VSMALL
constexpr double PI = 3.14159265358979323846;
constexpr int FK_INTERNAL = 0, FK_WALL = 1, FK_CYCLIC = 2;
constexpr int FK_FIXED = 3;   // frozen incoming Maxwellian rho_b M(U_b, T_b) (rho fixedValue, gSurf mixed)
constexpr int FK_ZG = 4;      // zeroGradient: every face value is the owner cell's gBarPvol
constexpr int FK_SYM = 5;     // symmetry plane: incoming value = outgoing value of the mirrored velocity
constexpr int FK_FAR = 6;     // far field: incoming rho_b M(U of the owner cell, T_b)
constexpr int FK_PIN = 7;     // pressureIn: a far field with rho_b = p_b / (R T_b)
constexpr int FK_POUT = 8;    // pressureOut: a far field whose rho_b, T_b follow the owner cell
// symmetryPlane (OpenFOAM constraint patch): incoming value as FK_SYM, the gradient takes
// phi_b = phi_P (no lagged row), S7 relaxes every value once and the outgoing ones once more
constexpr int FK_SPL = 9;

// Non-coupled boundary faces: a lagged fixedGradient row, only the outgoing half relaxed
// (FK_SPL is the exception to both).
ADJ_HD bool open_kind(int fk) { return fk != FK_INTERNAL && fk != FK_CYCLIC; }
// Boundary faces whose incoming half is a Maxwellian at the boundary temperature (dJ/dT output).
ADJ_HD bool maxwell_kind(int fk) { return fk == FK_WALL || fk == FK_FIXED || fk == FK_FAR || fk == FK_PIN || fk == FK_POUT; }
// Far-field kinds: incoming rho_b M(U of the owner cell, T_b).
ADJ_HD bool far_kind(int fk) { return fk == FK_FAR || fk == FK_PIN || fk == FK_POUT; }
// Mirror planes: the incoming value of a velocity is the outgoing value of its mirror.
ADJ_HD bool mirror_kind(int fk) { return fk == FK_SYM || fk == FK_SPL; }

// Mesh, velocity set and gas; every array is indexed by global face / cell numbers.
struct AP {
    int D, K, nc, nfi, nf, nv;
    double R, omega, Tref, mu, Pr, kc, dt;   // kc = K + 3 - D
    const int *cstart, *cface;   // faces of each cell in ascending order
    const int *owner, *nei;      // nei: internal faces
    const int *fkind;            // FK_*
    const int *partner, *pcell;  // cyclic faces: the coupled face and its owner cell
    const double *S, *Cf, *C, *V, *wf;   // wf: surfaceInterpolation weights
    const double *pvec, *nvec;   // least-squares vectors [nf][3], [nfi][3]
    const double *nhat, *adel;   // walls: Sf/|Sf|, 1/deltaCoeffs
    const double *xi, *wv;       // [nv][3], [nv], reference order
    const double *Tw, *Uw;       // walls, fixed and far-field faces: T [nf]; walls, fixed: U [nf][3]
    const double *rb;            // fixed and far-field faces: rho_b [nf]
    const double *win, *dwin;    // walls: inComingByRho and its T_w derivative
    const double *pb;            // pressure faces: p_b [nf]
    const double *muc;           // muRef of each cell's case [nc] (faces: the owner's)
    const int *sax;              // symmetry faces: the mirror axis [nf]
    const int *mir;              // [3][nv]: the velocity mirrored across axis a
};

// One state: g, h [nc][nv]; Gg, Gh [nf - nfi][nv]; mac [nc][8] = rho, U[3], T, q[3].
struct St {
    double *g, *h, *Gg, *Gh, *mac;
};

// Face quantities of S3/S6 kept for the reverse sweep.
struct FaceRec {
    double rho, U[3], T, q[3], tau, r, F, qh[3], A0, P[6], rho_w, pad[10];
};
static_assert(sizeof(FaceRec) == 32 * sizeof(double), "face record");

// xi . S = (x Sx + y Sy) + z Sz without contraction (the upwind test must see exact zeros)
ADJ_HD double sdot(const double *a, const double *b)
{
#pragma clang fp contract(off)
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

ADJ_HD double tau_of(const AP &p, double mu, double rho, double T)
{
    return mu * exp(p.omega * log(T / p.Tref)) / rho / T / p.R;   // fvDVM.C:808-817
}

struct Eq {
    double rho, U[3], T, q[3], RT, pre, k5;   // pre = rho / (2 pi R T)^(D/2), k5 = 1 / (5 rho (RT)^2)
};

ADJ_HD void eq_set(const AP &p, double rho, const double *U, double T, const double *q, Eq &e)
{
    e.rho = rho;
    e.T = T;
    for (int d = 0; d < 3; ++d) {
        e.U[d] = U[d];
        e.q[d] = q[d];
    }
    e.RT = p.R * T;
    e.pre = rho / pow(sqrt(2.0 * PI * e.RT), (double)p.D);
    e.k5 = 1.0 / (5.0 * rho * e.RT * e.RT);
}

// Shakhov equilibrium, discreteVelocity.C:1016-1044
ADJ_HD void shakhov(const AP &p, const Eq &e, const double *xi, double &geq, double &heq)
{
    const double u[3] = {e.U[0] - xi[0], e.U[1] - xi[1], e.U[2] - xi[2]};
    const double c2 = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / e.RT;
    const double cq = -(u[0] * e.q[0] + u[1] * e.q[1] + u[2] * e.q[2]) * e.k5;
    const double gB = e.pre * exp(-0.5 * c2);
    const double P1 = 1.0 - p.Pr;
    geq = (1.0 + P1 * cq * (c2 - p.D - 2.0)) * gB;
    heq = (p.kc + P1 * cq * ((c2 - p.D) * p.kc - 2.0 * p.K)) * gB * e.RT;
}

// acc[0..8) += a d(geq)/d(rho, U, T, q) + b d(heq)/d(rho, U, T, q)
ADJ_HD void shakhov_vjp(const AP &p, const Eq &e, const double *xi, double a, double b, double *acc)
{
    const double u[3] = {e.U[0] - xi[0], e.U[1] - xi[1], e.U[2] - xi[2]};
    const double RT = e.RT, T = e.T, rho = e.rho;
    const double c2 = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / RT;
    const double cq = -(u[0] * e.q[0] + u[1] * e.q[1] + u[2] * e.q[2]) * e.k5;
    const double gB = e.pre * exp(-0.5 * c2);
    const double P1 = 1.0 - p.Pr;
    const double D = p.D, kc = p.kc, K2 = 2.0 * p.K;
    const double Ag = 1.0 + P1 * cq * (c2 - D - 2.0);
    const double Ah = kc + P1 * cq * ((c2 - D) * kc - K2);
    const double X = a * Ag + b * RT * Ah;                       // multiplies d(gB)
    const double Y1 = a * (c2 - D - 2.0) + b * RT * ((c2 - D) * kc - K2);   // multiplies d(cq)
    const double Y2 = a + b * RT * kc;                           // multiplies cq d(c2)
    const double gp = gB * P1;
    acc[0] += gp * (-cq / rho) * Y1 + X * gB / rho;
    for (int d = 0; d < 3; ++d)
        acc[1 + d] += gp * (-e.q[d] * e.k5 * Y1 + cq * (2.0 * u[d] / RT) * Y2) - X * gB * u[d] / RT;
    acc[4] += gp * ((-2.0 * cq / T) * Y1 - cq * (c2 / T) * Y2) + X * gB * (c2 - D) / (2.0 * T) + b * Ah * gB * p.R;
    for (int d = 0; d < 3; ++d) acc[5 + d] += gp * (-u[d] * e.k5) * Y1;
}

// equilibriumMaxwellByRho (discreteVelocity.C:1063-1075) and its T derivative
ADJ_HD double maxwell(const AP &p, const double *U, double T, const double *xi, double &dMdT)
{
    const double u[3] = {U[0] - xi[0], U[1] - xi[1], U[2] - xi[2]};
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double M = 1.0 / pow(sqrt(2.0 * PI * p.R * T), (double)p.D) * exp(-u2 / (2.0 * p.R * T));
    dMdT = M * (-0.5 * p.D / T + u2 / (2.0 * p.R * T * T));
    return M;
}

// rho_b and T_b of a far-field face in the step that starts from the state with cell macros mac.
// Pressure patches take them from S10 of the previous step (fvDVM.C:730-806): inlet rho_b =
// p_b / (R T_b); outlet rho_b = rho_i + (p_b - rho_i R T_i) / a_i^2, T_b = p_b / (R rho_i) with
// a_i^2 = R T_i (K + 5) / (K + 3) of the owner cell i.  The first step takes them from the initial
// state, as the constructor runs S10 once before the first step (fvDVM.C:1071-1072).
ADJ_HD void far_bc(const AP &p, int f, int fk, const double *mac, double &rb, double &Tb)
{
    Tb = p.Tw[f];
    if (fk == FK_FAR) {
        rb = p.rb[f];
        return;
    }
    if (fk == FK_PIN) {
        rb = p.pb[f] / p.R / Tb;
        return;
    }
    const double *mi = mac + 8 * (size_t)p.owner[f];
    const double rhoi = mi[0], Ti = mi[4];
    const double ai = sqrt(p.R * Ti * (p.K + 5) / (p.K + 3));
    rb = rhoi + (p.pb[f] - rhoi * p.R * Ti) / ai / ai;
    Tb = p.pb[f] / (p.R * rhoi);
}

// Moments of one point about its velocity U: A0 = 1/2 sum w (|c|^2 g + h),
// P_ij = 1/2 sum w c_i c_j g, qh = 1/2 sum w c (|c|^2 g + h), c = xi - U.
struct Mom {
    double rho, U[3], T, tau, F, c0, qh[3], A0, P[3][3];
};

ADJ_HD void mom_acc(double *s, const double *xi, double w, const double *U, double g, double h)
{
    const double c[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
    const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double a = 0.5 * w * (cc * g + h);
    s[0] += c[0] * a;
    s[1] += c[1] * a;
    s[2] += c[2] * a;
    s[3] += a;
    const double hg = 0.5 * w * g;
    s[4] += hg * c[0] * c[0];
    s[5] += hg * c[0] * c[1];
    s[6] += hg * c[0] * c[2];
    s[7] += hg * c[1] * c[1];
    s[8] += hg * c[1] * c[2];
    s[9] += hg * c[2] * c[2];
}

ADJ_HD void mom_set(Mom &m, const double *s)
{
    for (int d = 0; d < 3; ++d) m.qh[d] = s[d];
    m.A0 = s[3];
    m.P[0][0] = s[4]; m.P[0][1] = m.P[1][0] = s[5]; m.P[0][2] = m.P[2][0] = s[6];
    m.P[1][1] = s[7]; m.P[1][2] = m.P[2][1] = s[8]; m.P[2][2] = s[9];
}

// Adjoint of the macro map (g_k, h_k) -> (rho, U, T, q) of fvDVM.C:473-522 / 612-727, with
// q = F(tau) qh, F = 2 tau / (2 tau + c0), tau = tau(rho, T).  ad[8] is the adjoint of
// (rho, U, T, q), tau_ext an extra adjoint of tau.  The adjoint of the distributions is then
//   g_k += w_k (al + be . xi + ga |xi|^2 + (de . c) |c|^2),  h_k += w_k (ga + de . c),  c = xi - U.
ADJ_HD void macro_adj(const AP &p, const Mom &m, const double *ad, double tau_ext, double &al, double *be,
                      double &ga, double *de)
{
    double Fh = 0.0;
    for (int d = 0; d < 3; ++d) Fh += ad[5 + d] * m.qh[d];
    const double t2c = 2.0 * m.tau + m.c0;
    const double th = tau_ext + Fh * 2.0 * m.c0 / (t2c * t2c);
    double qa[3], Ut[3];
    for (int d = 0; d < 3; ++d) {
        qa[d] = m.F * ad[5 + d];
        de[d] = 0.5 * qa[d];
    }
    for (int e = 0; e < 3; ++e) {
        double s = 0.0;
        for (int d = 0; d < 3; ++d) s += qa[d] * m.P[d][e];
        Ut[e] = ad[1 + e] - qa[e] * m.A0 - 2.0 * s;
    }
    const double Tt = ad[4] + th * m.tau * (p.omega - 1.0) / m.T;
    const double r1 = ad[0] - th * m.tau / m.rho;
    const double den = 0.5 * (p.K + 3) * p.R * m.rho;
    double uu = 0.0, uU = 0.0;
    for (int d = 0; d < 3; ++d) {
        be[d] = Ut[d] / m.rho - Tt * m.U[d] / den;
        uu += m.U[d] * m.U[d];
        uU += Ut[d] * m.U[d];
    }
    al = r1 - uU / m.rho + Tt * (0.5 * uu / den - m.T / m.rho);
    ga = 0.5 * Tt / den;
}

ADJ_HD void macro_apply(const double *xi, double w, const double *U, double al, const double *be, double ga,
                        const double *de, double &ag, double &ah)
{
    const double c[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
    const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    const double xx = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    const double dc = de[0] * c[0] + de[1] * c[1] + de[2] * c[2];
    ag += w * (al + (be[0] * xi[0] + be[1] * xi[1] + be[2] * xi[2]) + ga * xx + dc * cc);
    ah += w * (ga + dc);
}

ADJ_HD size_t ix(int point, int nv, int k) { return (size_t)point * nv + k; }

// grad layout [cell][d < D][k]
ADJ_HD void load_grad(const AP &p, const double *G, int c, int k, double *v)
{
    for (int d = 0; d < 3; ++d) v[d] = d < p.D ? G[((size_t)c * p.D + d) * p.nv + k] : 0.0;
}

ADJ_HD double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ---------------------------------------------------------------------------------------------
// Forward step
// ---------------------------------------------------------------------------------------------

// initDFtoEq (discreteVelocity.C:220-249): Shakhov with q = 0; macros = the given fields
template <class X>
ADJ_HD void init_cell(X &x, const AP &p, St s, int c, const double *rho0, const double *U0, const double *T0)
{
    const double q0[3] = {0, 0, 0};
    Eq e;
    eq_set(p, rho0[c], U0 + 3 * c, T0[c], q0, e);
    x.for_k(p.nv, [&](int k) {
        double g, h;
        shakhov(p, e, p.xi + 3 * k, g, h);
        s.g[ix(c, p.nv, k)] = g;
        s.h[ix(c, p.nv, k)] = h;
    });
    if (x.leader()) {
        double *m = s.mac + 8 * (size_t)c;
        m[0] = rho0[c];
        for (int d = 0; d < 3; ++d) {
            m[1 + d] = U0[3 * c + d];
            m[5 + d] = 0.0;
        }
        m[4] = T0[c];
    }
}

// inComingByRho of one wall face (fvDVM.C:263-309) and its derivative in T_w
template <class X>
ADJ_HD void wall_in(X &x, const AP &p, int f, double *win, double *dwin)
{
    double a[2] = {0.0, 0.0};
    const double *S = p.S + 3 * f;
    x.for_k(p.nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const double s = sdot(xi, S);
        if (s < 0) {
            double dM;
            const double M = maxwell(p, p.Uw + 3 * f, p.Tw[f], xi, dM);
            a[0] += -p.wv[k] * s * M;
            a[1] += -p.wv[k] * s * dM;
        }
    });
    x.template sum<2>(a);
    if (x.leader()) {
        win[f] = a[0];
        dwin[f] = a[1];
    }
}

ADJ_HD void cell_macros(const AP &p, const double *m, double &rho, double *U, double &T, double *q)
{
    rho = m[0];
    for (int d = 0; d < 3; ++d) {
        U[d] = m[1 + d];
        q[d] = m[5 + d];
    }
    T = m[4];
}

// S1 (discreteVelocity.C:393-406): gbar+ = (1 - r) g~ + r geq, r = 1.5 dt / (2 tau + dt)
template <class X>
ADJ_HD void s1_cell(X &x, const AP &p, St s, int c, double *gp, double *hp)
{
    double rho, U[3], T, q[3];
    cell_macros(p, s.mac + 8 * (size_t)c, rho, U, T, q);
    const double tau = tau_of(p, p.muc[c], rho, T);
    const double r = 1.5 * p.dt / (2.0 * tau + p.dt);
    Eq e;
    eq_set(p, rho, U, T, q, e);
    x.for_k(p.nv, [&](int k) {
        double geq, heq;
        shakhov(p, e, p.xi + 3 * k, geq, heq);
        const size_t i = ix(c, p.nv, k);
        gp[i] = (1.0 - r) * s.g[i] + r * geq;
        hp[i] = (1.0 - r) * s.h[i] + r * heq;
    });
}

// S2 least-squares gradient of gbar+ (leastSquaresGrad with the lagged fixedGradient wall rows
// phi_b = phi_P + G / deltaCoeffs) and, when Gn is given, the next step's rows n . grad
// (discreteVelocity.C:412-470)
template <class X>
ADJ_HD void grad_cell(X &x, const AP &p, St s, int c, const double *gp, const double *hp, double *gg, double *gh,
                      double *Gng, double *Gnh)
{
    const int nfi = p.nfi, nv = p.nv;
    x.for_k(nv, [&](int k) {
        const double gc = gp[ix(c, nv, k)], hc = hp[ix(c, nv, k)];
        double G[3] = {0, 0, 0}, H[3] = {0, 0, 0};
        for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
            const int f = p.cface[e];
            const double *L;
            double dg, dh;
            if (f < nfi) {
                const bool own = p.owner[f] == c;
                const int o = own ? p.nei[f] : p.owner[f];
                L = own ? p.pvec + 3 * f : p.nvec + 3 * f;
                dg = gp[ix(o, nv, k)] - gc;
                dh = hp[ix(o, nv, k)] - hc;
            } else if (p.fkind[f] == FK_CYCLIC) {
                const int o = p.pcell[f];
                L = p.pvec + 3 * f;
                dg = gp[ix(o, nv, k)] - gc;
                dh = hp[ix(o, nv, k)] - hc;
            } else if (p.fkind[f] == FK_SPL) {
                L = p.pvec + 3 * f;
                dg = gc - gc;   // phi_b = phi_P
                dh = hc - hc;
            } else {
                L = p.pvec + 3 * f;
                const size_t b = ix(f - nfi, nv, k);
                dg = (gc + s.Gg[b] * p.adel[f]) - gc;
                dh = (hc + s.Gh[b] * p.adel[f]) - hc;
            }
            for (int d = 0; d < 3; ++d) {
                G[d] += L[d] * dg;
                H[d] += L[d] * dh;
            }
        }
        for (int d = 0; d < p.D; ++d) {
            gg[((size_t)c * p.D + d) * nv + k] = G[d];
            gh[((size_t)c * p.D + d) * nv + k] = H[d];
        }
        if (Gng)
            for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
                const int f = p.cface[e];
                if (f < nfi || p.fkind[f] == FK_CYCLIC || p.fkind[f] == FK_SPL) continue;
                Gng[ix(f - nfi, nv, k)] = dot3(G, p.nhat + 3 * f);
                Gnh[ix(f - nfi, nv, k)] = dot3(H, p.nhat + 3 * f);
            }
    });
}

// upwind face value (discreteVelocity.C:472-530 internal, 605-627 wall outgoing, 628-672 cyclic)
ADJ_HD void face_value(const AP &p, int f, int k, double s, const double *gp, const double *hp, const double *gg,
                       const double *gh, double &g, double &h)
{
    const int nv = p.nv, P = p.owner[f], fk = f < p.nfi ? FK_INTERNAL : p.fkind[f];
    const double *xi = p.xi + 3 * k, *Cf = p.Cf + 3 * f;
    const double hx[3] = {0.5 * xi[0] * p.dt, 0.5 * xi[1] * p.dt, 0.5 * xi[2] * p.dt};
    double GP[3], HP[3];
    load_grad(p, gg, P, k, GP);
    load_grad(p, gh, P, k, HP);
    const double *CP = p.C + 3 * P;
    const double dO[3] = {Cf[0] - CP[0] - hx[0], Cf[1] - CP[1] - hx[1], Cf[2] - CP[2] - hx[2]};
    if (fk == FK_ZG) {
        g = gp[ix(P, nv, k)];   // patchInternalField (discreteVelocity.C:551-555)
        h = hp[ix(P, nv, k)];
        return;
    }
    if (mirror_kind(fk) && !(s > 0)) {
        // incoming: the outgoing reconstruction of the mirrored velocity, at its own index
        // (discreteVelocity.C:746-790); xi.S == 0 mirrors onto itself
        const int km = p.mir[p.sax[f] * nv + k];
        const double *xm = p.xi + 3 * km;
        const double hm[3] = {0.5 * xm[0] * p.dt, 0.5 * xm[1] * p.dt, 0.5 * xm[2] * p.dt};
        double GM[3], HM[3];
        load_grad(p, gg, P, km, GM);
        load_grad(p, gh, P, km, HM);
        const double dM[3] = {Cf[0] - CP[0] - hm[0], Cf[1] - CP[1] - hm[1], Cf[2] - CP[2] - hm[2]};
        g = gp[ix(P, nv, km)] + dot3(GM, dM);
        h = hp[ix(P, nv, km)] + dot3(HM, dM);
        return;
    }
    const double gP = gp[ix(P, nv, k)] + dot3(GP, dO), hP = hp[ix(P, nv, k)] + dot3(HP, dO);
    if (fk == FK_WALL || fk == FK_FIXED || mirror_kind(fk) || far_kind(fk)) {
        g = gP;
        h = hP;
        return;
    }
    if (fk == FK_INTERNAL) {
        if (s >= VSMALL) {
            g = gP;
            h = hP;
            return;
        }
        const int N = p.nei[f];
        double GN[3], HN[3];
        load_grad(p, gg, N, k, GN);
        load_grad(p, gh, N, k, HN);
        const double *CN = p.C + 3 * N;
        const double dN[3] = {Cf[0] - CN[0] - hx[0], Cf[1] - CN[1] - hx[1], Cf[2] - CN[2] - hx[2]};
        const double gN = gp[ix(N, nv, k)] + dot3(GN, dN), hN = hp[ix(N, nv, k)] + dot3(HN, dN);
        if (s < -VSMALL) {
            g = gN;
            h = hN;
        } else {
            g = 0.5 * (gN + gP);
            h = 0.5 * (hN + hP);
        }
        return;
    }
    // cyclic: coupled value w phi_P + (1 - w) phi_N and gradient, at distance -hx
    if (s > VSMALL) {
        g = gP;
        h = hP;
        return;
    }
    const int N = p.pcell[f];
    const double w = p.wf[f];
    double GN[3], HN[3];
    load_grad(p, gg, N, k, GN);
    load_grad(p, gh, N, k, HN);
    double gb = w * gp[ix(P, nv, k)] + (1.0 - w) * gp[ix(N, nv, k)];
    double hb = w * hp[ix(P, nv, k)] + (1.0 - w) * hp[ix(N, nv, k)];
    for (int d = 0; d < 3; ++d) {
        gb += (w * GP[d] + (1.0 - w) * GN[d]) * (-hx[d]);
        hb += (w * HP[d] + (1.0 - w) * HN[d]) * (-hx[d]);
    }
    if (s < -VSMALL) {
        g = gb;
        h = hb;
    } else {
        g = 0.5 * (gP + gb);
        h = 0.5 * (hP + hb);
    }
}

// S2-S7 of one face: face values (S2), Maxwell wall density and incoming values (S3, S4), face
// macros (S6) and, when gr is given, the relaxed face values (S7).  jacc[f] += c_f (rho U)_f . S_f
// when jacc is given.
template <class X>
ADJ_HD void face_fwd(X &x, const AP &p, int f, const double *mac, const double *gp, const double *hp,
                     const double *gg, const double *gh, double *gf, double *hf, FaceRec *fr, double *gr,
                     double *hr, const double *coef, double *jacc)
{
    const int nv = p.nv, fk = f < p.nfi ? FK_INTERNAL : p.fkind[f];
    const double *S = p.S + 3 * f;
    double out[1] = {0.0};
    x.for_k(nv, [&](int k) {
        const double s = sdot(p.xi + 3 * k, S);
        if (maxwell_kind(fk) && !(s > 0)) return;
        double g, h;
        face_value(p, f, k, s, gp, hp, gg, gh, g, h);
        gf[ix(f, nv, k)] = g;
        hf[ix(f, nv, k)] = h;
        if (fk == FK_WALL) out[0] += p.wv[k] * s * g;
    });
    double rho_w = 0.0;
    if (fk == FK_WALL) {
        x.template sum<1>(out);
        rho_w = out[0] / p.win[f];   // calculatedMaxwell evaluate (cMaxwell.C:139-164)
        const double Tw = p.Tw[f];
        x.for_k(nv, [&](int k) {
            const double *xi = p.xi + 3 * k;
            if (sdot(xi, S) > 0) return;
            double dM;
            const double g = rho_w * maxwell(p, p.Uw + 3 * f, Tw, xi, dM);
            gf[ix(f, nv, k)] = g;
            hf[ix(f, nv, k)] = g * (p.R * Tw) * p.kc;
        });
    }
    if (fk == FK_FIXED || far_kind(fk)) {
        // incoming rho_b M(U, T_b): fixed faces freeze U_b (discreteVelocity.C:312-344, 556-573),
        // far-field faces take the owner cell's velocity (discreteVelocity.C:588-602)
        double Tb = p.Tw[f], rb = p.rb[f];
        if (far_kind(fk)) far_bc(p, f, fk, mac, rb, Tb);
        const double *Ub = far_kind(fk) ? mac + 8 * (size_t)p.owner[f] + 1 : p.Uw + 3 * f;
        x.for_k(nv, [&](int k) {
            const double *xi = p.xi + 3 * k;
            if (sdot(xi, S) > 0) return;
            double dM;
            const double g = rb * maxwell(p, Ub, Tb, xi, dM);
            gf[ix(f, nv, k)] = g;
            hf[ix(f, nv, k)] = g * (p.R * Tb) * p.kc;
        });
    }
    double a[5] = {0, 0, 0, 0, 0};
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const double w = p.wv[k], g = gf[ix(f, nv, k)], h = hf[ix(f, nv, k)];
        a[0] += w * g;
        for (int d = 0; d < 3; ++d) a[1 + d] += w * g * xi[d];
        a[4] += 0.5 * w * (g * dot3(xi, xi) + h);
    });
    x.template sum<5>(a);
    Mom m;
    m.rho = a[0];
    for (int d = 0; d < 3; ++d) m.U[d] = a[1 + d] / m.rho;
    m.T = (a[4] - 0.5 * m.rho * dot3(m.U, m.U)) / ((p.K + 3) / 2.0 * p.R * m.rho);
    m.tau = tau_of(p, p.muc[p.owner[f]], m.rho, m.T);
    double sm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    x.for_k(nv, [&](int k) { mom_acc(sm, p.xi + 3 * k, p.wv[k], m.U, gf[ix(f, nv, k)], hf[ix(f, nv, k)]); });
    x.template sum<10>(sm);
    mom_set(m, sm);
    const double hdt = 0.5 * p.dt;
    m.c0 = 0.5 * p.dt * p.Pr;
    m.F = 2.0 * m.tau / (2.0 * m.tau + m.c0);   // fvDVM.C:522
    double q[3];
    for (int d = 0; d < 3; ++d) q[d] = m.F * m.qh[d];
    const double r = hdt / (2.0 * m.tau + hdt);   // discreteVelocity.C:867
    if (x.leader()) {
        FaceRec &R = fr[f];
        R.rho = m.rho;
        R.T = m.T;
        R.tau = m.tau;
        R.r = r;
        R.F = m.F;
        R.A0 = m.A0;
        R.rho_w = rho_w;
        for (int d = 0; d < 3; ++d) {
            R.U[d] = m.U[d];
            R.q[d] = q[d];
            R.qh[d] = m.qh[d];
        }
        for (int i = 0; i < 6; ++i) R.P[i] = sm[4 + i];
        if (jacc && coef[f] != 0.0) jacc[f] += coef[f] * (a[1] * S[0] + a[2] * S[1] + a[3] * S[2]);
    }
    if (!gr) return;
    Eq e;
    eq_set(p, m.rho, m.U, m.T, q, e);
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const size_t i = ix(f, nv, k);
        const bool out = sdot(xi, S) > 0;
        if (open_kind(fk) && fk != FK_SPL && !out) {
            gr[i] = gf[i];   // incoming halves are not relaxed (their patch operators are no-ops, *Fvs.H:161-181)
            hr[i] = hf[i];
            return;
        }
        double geq, heq;
        shakhov(p, e, xi, geq, heq);
        gr[i] = (1.0 - r) * gf[i] + r * geq;
        hr[i] = (1.0 - r) * hf[i] + r * heq;
        if (fk == FK_SPL && out) {
            // symmetryPlane: the field operation relaxed every value (discreteVelocity.C:880-881),
            // the loop over non-coupled patches relaxes the outgoing ones again (905-920)
            gr[i] = (1.0 - r) * gr[i] + r * geq;
            hr[i] = (1.0 - r) * hr[i] + r * heq;
        }
    });
}

// S8 (discreteVelocity.C:934-978) and S9 (fvDVM.C:597-728) of one cell into state nx
template <class X>
ADJ_HD void cell_fwd(X &x, const AP &p, St s, int c, const double *gp, const double *hp, const double *gr,
                     const double *hr, St nx)
{
    const int nv = p.nv;
    const double V = p.V[c], dt = p.dt;
    double a[5] = {0, 0, 0, 0, 0};
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const size_t i = ix(c, nv, k);
        double g = -1.0 / 3 * s.g[i] + 4.0 / 3 * gp[i];
        double h = -1.0 / 3 * s.h[i] + 4.0 / 3 * hp[i];
        for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
            const int f = p.cface[e];
            const double sf = sdot(xi, p.S + 3 * f);
            const double fg = sf * gr[ix(f, nv, k)] * dt / V, fh = sf * hr[ix(f, nv, k)] * dt / V;
            if (p.owner[f] == c) {
                g -= fg;
                h -= fh;
            } else {
                g += fg;
                h += fh;
            }
        }
        nx.g[i] = g;
        nx.h[i] = h;
        const double w = p.wv[k];
        a[0] += w * g;
        for (int d = 0; d < 3; ++d) a[1 + d] += w * g * xi[d];
        a[4] += 0.5 * w * (dot3(xi, xi) * g + h);
    });
    x.template sum<5>(a);
    const double rho = a[0];
    double U[3];
    for (int d = 0; d < 3; ++d) U[d] = a[1 + d] / rho;
    const double T = (a[4] - 0.5 * rho * dot3(U, U)) / ((p.K + 3) / 2.0 * p.R * rho);
    const double tau = tau_of(p, p.muc[c], rho, T);
    double sm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    x.for_k(nv, [&](int k) { mom_acc(sm, p.xi + 3 * k, p.wv[k], U, nx.g[ix(c, nv, k)], nx.h[ix(c, nv, k)]); });
    x.template sum<10>(sm);
    const double F = 2.0 * tau / (2.0 * tau + dt * p.Pr);   // fvDVM.C:727
    if (x.leader()) {
        double *m = nx.mac + 8 * (size_t)c;
        m[0] = rho;
        for (int d = 0; d < 3; ++d) {
            m[1 + d] = U[d];
            m[5 + d] = F * sm[d];
        }
        m[4] = T;
    }
}

// ---------------------------------------------------------------------------------------------
// Reverse step
// ---------------------------------------------------------------------------------------------

// S8 -> S7 -> objective -> S6 -> S4/S3 of one face.  ag1, ah1: adjoint of g~^{n+1}, h~^{n+1}.
// Writes the adjoint of the reconstructed face values (zero for incoming wall values) and adds
// this step's dJ/dT_w of a wall face.  coef null: the step is outside the objective's window.
template <class X>
ADJ_HD void face_adj(X &x, const AP &p, int f, const double *mac, const FaceRec *fr, const double *gf,
                     const double *hf, const double *ag1, const double *ah1, const double *coef, double *abf,
                     double *abh, double *dTw, double *aUf)
{
    const int nv = p.nv, fk = f < p.nfi ? FK_INTERNAL : p.fkind[f];
    const double *S = p.S + 3 * f;
    const FaceRec &R = fr[f];
    const int P = p.owner[f], N = fk == FK_INTERNAL ? p.nei[f] : -1;
    const double dt = p.dt, VP = p.V[P], VN = N >= 0 ? p.V[N] : 1.0;
    const double cf = coef ? coef[f] : 0.0;
    Eq e;
    eq_set(p, R.rho, R.U, R.T, R.q, e);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const double s = sdot(xi, S);
        const size_t i = ix(f, nv, k);
        double af = -(s * dt) * ag1[ix(P, nv, k)] / VP, bf = -(s * dt) * ah1[ix(P, nv, k)] / VP;
        if (N >= 0) {
            af += (s * dt) * ag1[ix(N, nv, k)] / VN;
            bf += (s * dt) * ah1[ix(N, nv, k)] / VN;
        }
        double ab = af, bb = bf;
        if (fk == FK_SPL && s > 0) {
            // two relaxations, g1 = (1 - r) gf + r geq and gr = (1 - r) g1 + r geq
            double geq, heq;
            shakhov(p, e, xi, geq, heq);
            const double g1 = (1.0 - R.r) * gf[i] + R.r * geq, h1 = (1.0 - R.r) * hf[i] + R.r * heq;
            acc[0] += af * (geq - g1) + bf * (heq - h1) + (1.0 - R.r) * (af * (geq - gf[i]) + bf * (heq - hf[i]));
            const double we = R.r * (2.0 - R.r);
            shakhov_vjp(p, e, xi, we * af, we * bf, acc + 1);
            ab = (1.0 - R.r) * (1.0 - R.r) * af;
            bb = (1.0 - R.r) * (1.0 - R.r) * bf;
        } else if (!open_kind(fk) || fk == FK_SPL || s > 0) {
            double geq, heq;
            shakhov(p, e, xi, geq, heq);
            acc[0] += af * (geq - gf[i]) + bf * (heq - hf[i]);
            shakhov_vjp(p, e, xi, R.r * af, R.r * bf, acc + 1);
            ab = (1.0 - R.r) * af;
            bb = (1.0 - R.r) * bf;
        }
        if (cf != 0.0) ab += cf * p.wv[k] * s;
        abf[i] = ab;
        abh[i] = bb;
    });
    x.template sum<9>(acc);
    const double hdt = 0.5 * dt;
    const double th = acc[0] * (-2.0 * R.r / (2.0 * R.tau + hdt));
    Mom m;
    m.rho = R.rho;
    m.T = R.T;
    m.tau = R.tau;
    m.F = R.F;
    m.c0 = 0.5 * dt * p.Pr;
    m.A0 = R.A0;
    for (int d = 0; d < 3; ++d) {
        m.U[d] = R.U[d];
        m.qh[d] = R.qh[d];
    }
    {
        double sm[10];
        for (int i = 0; i < 6; ++i) sm[4 + i] = R.P[i];
        for (int d = 0; d < 3; ++d) sm[d] = R.qh[d];
        sm[3] = R.A0;
        mom_set(m, sm);
    }
    double al, be[3], ga, de[3];
    macro_adj(p, m, acc + 1, th, al, be, ga, de);
    double Tw = maxwell_kind(fk) ? p.Tw[f] : 0.0;
    double rbf = fk == FK_FIXED || fk == FK_FAR ? p.rb[f] : 0.0;
    if (far_kind(fk)) far_bc(p, f, fk, mac, rbf, Tw);
    const double hfac = p.R * Tw * p.kc;
    const double *Ub = far_kind(fk) ? mac + 8 * (size_t)p.owner[f] + 1 : p.Uw + 3 * f;
    // walls: M sum, dT; fixed and far-field kinds: dM sum (d rho_b), dT_b, dU of the owner (far)
    double wa[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const size_t i = ix(f, nv, k);
        double ab = abf[i], bb = abh[i];
        macro_apply(xi, p.wv[k], m.U, al, be, ga, de, ab, bb);
        if (fk == FK_WALL && !(sdot(xi, S) > 0)) {
            double dM;
            const double M = maxwell(p, p.Uw + 3 * f, Tw, xi, dM);
            const double t = ab + bb * hfac;
            wa[0] += t * M;
            wa[1] += t * R.rho_w * dM + bb * R.rho_w * M * p.R * p.kc;
            ab = 0.0;
            bb = 0.0;
        } else if ((fk == FK_FIXED || far_kind(fk)) && !(sdot(xi, S) > 0)) {
            // incoming rho_b M(U, T_b): the rho_b and T_b derivatives, and on far-field faces the
            // derivative in the owner cell's velocity, dM/dU = M (xi - U) / (R T_b)
            double dM;
            const double M = maxwell(p, Ub, Tw, xi, dM);
            const double t = ab + bb * hfac;
            wa[0] += t * M;
            wa[1] += t * rbf * dM + bb * rbf * M * p.R * p.kc;
            if (far_kind(fk))
                for (int d = 0; d < 3; ++d) wa[2 + d] += t * rbf * M * (xi[d] - Ub[d]) / (p.R * Tw);
            ab = 0.0;
            bb = 0.0;
        }
        abf[i] = ab;
        abh[i] = bb;
    });
    if (fk == FK_FIXED || far_kind(fk)) {
        x.template sum<5>(wa);
        if (x.leader()) {
            // wa[0] = dJ/drho_b, wa[1] = dJ/dT_b at fixed rho_b, wa[2..4] = dJ/dU of the owner
            double dT = wa[1], drho_i = 0.0, dT_i = 0.0;
            if (fk == FK_PIN) dT += wa[0] * (-p.pb[f] / (p.R * Tw * Tw));
            if (fk == FK_POUT) {
                // S10 of the previous step: rho_b(rho_i, T_i), T_b(rho_i); the patch T is unused
                const double *mi = mac + 8 * (size_t)p.owner[f];
                const double rhoi = mi[0], Ti = mi[4], cK = p.R * (p.K + 5) / (p.K + 3);
                drho_i = wa[0] * (1.0 - p.R / cK) + wa[1] * (-p.pb[f] / (p.R * rhoi * rhoi));
                dT_i = wa[0] * (-p.pb[f] / (cK * Ti * Ti));
                dT = 0.0;
            }
            dTw[f] += dT;
            if (far_kind(fk)) {
                for (int d = 0; d < 3; ++d) aUf[5 * (size_t)f + d] = wa[2 + d];
                aUf[5 * (size_t)f + 3] = drho_i;
                aUf[5 * (size_t)f + 4] = dT_i;
            }
        }
        return;
    }
    if (fk != FK_WALL) return;
    x.template sum<2>(wa);
    const double oh = wa[0] / p.win[f], ih = -wa[0] * R.rho_w / p.win[f];
    if (x.leader()) dTw[f] += wa[1] + ih * p.dwin[f];
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const double s = sdot(xi, S);
        if (s > 0) abf[ix(f, nv, k)] += oh * p.wv[k] * s;
    });
}

// Reverse of the reconstruction gathered at one cell: the adjoint of gbar+ (with the 4/3 of S8)
// and of its gradient, which also takes the adjoint of the next lagged rows G^{n+1} = n . grad.
template <class X>
ADJ_HD void recon_adj(X &x, const AP &p, int c, const double *abf, const double *abh, const double *ag1,
                      const double *ah1, const double *AGg, const double *AGh, double *apg, double *aph,
                      double *agg, double *agh)
{
    const int nv = p.nv, nfi = p.nfi;
    const double *Cc = p.C + 3 * c;
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const double hx[3] = {0.5 * xi[0] * p.dt, 0.5 * xi[1] * p.dt, 0.5 * xi[2] * p.dt};
        const size_t i = ix(c, nv, k);
        double a = 4.0 / 3 * ag1[i], ah = 4.0 / 3 * ah1[i];
        double G[3] = {0, 0, 0}, H[3] = {0, 0, 0};
        for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
            const int f = p.cface[e];
            const double s = sdot(xi, p.S + 3 * f);
            const double *Cf = p.Cf + 3 * f;
            const double d[3] = {Cf[0] - Cc[0] - hx[0], Cf[1] - Cc[1] - hx[1], Cf[2] - Cc[2] - hx[2]};
            const double vg = abf[ix(f, nv, k)], vh = abh[ix(f, nv, k)];
            if (f < nfi) {
                const bool own = p.owner[f] == c;
                const double om = own ? (s >= VSMALL ? 1.0 : (s < -VSMALL ? 0.0 : 0.5))
                                      : (s < -VSMALL ? 1.0 : (s >= VSMALL ? 0.0 : 0.5));
                a += om * vg;
                ah += om * vh;
                for (int q = 0; q < 3; ++q) {
                    G[q] += om * vg * d[q];
                    H[q] += om * vh * d[q];
                }
            } else if (p.fkind[f] == FK_ZG) {
                a += vg;   // the cell's own value on every velocity
                ah += vh;
                const double bg = AGg[ix(f - nfi, nv, k)], bh = AGh[ix(f - nfi, nv, k)];
                for (int q = 0; q < 3; ++q) {
                    G[q] += bg * p.nhat[3 * f + q];
                    H[q] += bh * p.nhat[3 * f + q];
                }
            } else if (p.fkind[f] != FK_CYCLIC) {
                // outgoing value of this velocity; on a symmetry plane also the incoming value of
                // its mirror, which is the same reconstruction
                double wg = s > 0 ? vg : 0.0, wh = s > 0 ? vh : 0.0;
                if (mirror_kind(p.fkind[f])) {
                    const int km = p.mir[p.sax[f] * nv + k];
                    if (!(sdot(p.xi + 3 * km, p.S + 3 * f) > 0)) {
                        wg += abf[ix(f, nv, km)];
                        wh += abh[ix(f, nv, km)];
                    }
                }
                a += wg;
                ah += wh;
                for (int q = 0; q < 3; ++q) {
                    G[q] += wg * d[q];
                    H[q] += wh * d[q];
                }
                if (p.fkind[f] != FK_SPL) {   // a symmetryPlane has no lagged row
                    const double bg = AGg[ix(f - nfi, nv, k)], bh = AGh[ix(f - nfi, nv, k)];
                    for (int q = 0; q < 3; ++q) {
                        G[q] += bg * p.nhat[3 * f + q];
                        H[q] += bh * p.nhat[3 * f + q];
                    }
                }
            } else {
                // own face: upwind part (weight o1) and coupled part (o2), w = this face's weight
                const double w = p.wf[f];
                const double o1 = s > VSMALL ? 1.0 : (s < -VSMALL ? 0.0 : 0.5), o2 = 1.0 - o1;
                a += (o1 + o2 * w) * vg;
                ah += (o1 + o2 * w) * vh;
                for (int q = 0; q < 3; ++q) {
                    G[q] += o1 * vg * d[q] - o2 * w * vg * hx[q];
                    H[q] += o1 * vh * d[q] - o2 * w * vh * hx[q];
                }
                // the partner face sees this cell as its other side, weight 1 - w_partner
                const int fp = p.partner[f];
                const double sp = sdot(xi, p.S + 3 * fp);
                const double o3 = sp > VSMALL ? 0.0 : (sp < -VSMALL ? 1.0 : 0.5);
                const double wp = 1.0 - p.wf[fp];
                const double pg = abf[ix(fp, nv, k)], ph = abh[ix(fp, nv, k)];
                a += o3 * wp * pg;
                ah += o3 * wp * ph;
                for (int q = 0; q < 3; ++q) {
                    G[q] -= o3 * wp * pg * hx[q];
                    H[q] -= o3 * wp * ph * hx[q];
                }
            }
        }
        apg[i] = a;
        aph[i] = ah;
        for (int d = 0; d < p.D; ++d) {
            agg[((size_t)c * p.D + d) * nv + k] = G[d];
            agh[((size_t)c * p.D + d) * nv + k] = H[d];
        }
    });
}

// Reverse of the least-squares gradient gathered at one cell: adds to the adjoint of gbar+ and
// writes the adjoint of the lagged rows G^n of its wall faces.
template <class X>
ADJ_HD void grad_adj(X &x, const AP &p, int c, const double *agg, const double *agh, double *apg, double *aph,
                     double *AGg, double *AGh)
{
    const int nv = p.nv, nfi = p.nfi;
    x.for_k(nv, [&](int k) {
        const size_t i = ix(c, nv, k);
        double a = apg[i], ah = aph[i];
        double Gc[3], Hc[3];
        load_grad(p, agg, c, k, Gc);
        load_grad(p, agh, c, k, Hc);
        for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
            const int f = p.cface[e];
            const double *Lc, *Lo;
            int o;
            if (f < nfi) {
                const bool own = p.owner[f] == c;
                o = own ? p.nei[f] : p.owner[f];
                Lc = own ? p.pvec + 3 * f : p.nvec + 3 * f;
                Lo = own ? p.nvec + 3 * f : p.pvec + 3 * f;
            } else if (p.fkind[f] == FK_CYCLIC) {
                o = p.pcell[f];
                Lc = p.pvec + 3 * f;
                Lo = p.pvec + 3 * p.partner[f];
            } else {
                // lagged row: phi_b = phi_P + G / deltaCoeffs; a symmetryPlane's phi_b = phi_P
                // adds L (phi_P - phi_P) = 0 and reads no row
                const bool spl = p.fkind[f] == FK_SPL;
                AGg[ix(f - nfi, nv, k)] = spl ? 0.0 : p.adel[f] * dot3(p.pvec + 3 * f, Gc);
                AGh[ix(f - nfi, nv, k)] = spl ? 0.0 : p.adel[f] * dot3(p.pvec + 3 * f, Hc);
                continue;
            }
            double Go[3], Ho[3];
            load_grad(p, agg, o, k, Go);
            load_grad(p, agh, o, k, Ho);
            a += dot3(Lo, Go) - dot3(Lc, Gc);
            ah += dot3(Lo, Ho) - dot3(Lc, Hc);
        }
        apg[i] = a;
        aph[i] = ah;
    });
}

// Reverse of S1 at one cell, then the adjoint of the state's macros (S9 of the previous step)
// folded into the distributions: writes the adjoint of g~^n, h~^n.
template <class X>
ADJ_HD void s1_adj(X &x, const AP &p, St s, int c, const double *ag1, const double *ah1, const double *apg,
                   const double *aph, const double *aUf, double *ag0, double *ah0)
{
    const int nv = p.nv;
    double rho, U[3], T, q[3];
    cell_macros(p, s.mac + 8 * (size_t)c, rho, U, T, q);
    const double tau = tau_of(p, p.muc[c], rho, T);
    const double r = 1.5 * p.dt / (2.0 * tau + p.dt);
    Eq e;
    eq_set(p, rho, U, T, q, e);
    double acc[19];
    for (int i = 0; i < 19; ++i) acc[i] = 0.0;
    x.for_k(nv, [&](int k) {
        const double *xi = p.xi + 3 * k;
        const size_t i = ix(c, nv, k);
        const double a = apg[i], b = aph[i];
        double geq, heq;
        shakhov(p, e, xi, geq, heq);
        acc[0] += a * (geq - s.g[i]) + b * (heq - s.h[i]);
        shakhov_vjp(p, e, xi, r * a, r * b, acc + 1);
        mom_acc(acc + 9, xi, p.wv[k], U, s.g[i], s.h[i]);
    });
    x.template sum<19>(acc);
    // far-field faces of the cell: their incoming Maxwellians took this state's cell velocity
    // (and, on pressure outlets, rho_b and T_b from its density and temperature)
    for (int e = p.cstart[c]; e < p.cstart[c + 1]; ++e) {
        const int f = p.cface[e];
        if (f >= p.nfi && far_kind(p.fkind[f])) {
            acc[1] += aUf[5 * (size_t)f + 3];
            for (int d = 0; d < 3; ++d) acc[2 + d] += aUf[5 * (size_t)f + d];
            acc[5] += aUf[5 * (size_t)f + 4];
        }
    }
    const double th = acc[0] * (-2.0 * r / (2.0 * tau + p.dt));
    Mom m;
    m.rho = rho;
    m.T = T;
    m.tau = tau;
    m.c0 = p.dt * p.Pr;
    m.F = 2.0 * tau / (2.0 * tau + m.c0);
    for (int d = 0; d < 3; ++d) m.U[d] = U[d];
    mom_set(m, acc + 9);
    double al, be[3], ga, de[3];
    macro_adj(p, m, acc + 1, th, al, be, ga, de);
    x.for_k(nv, [&](int k) {
        const size_t i = ix(c, nv, k);
        double a = -1.0 / 3 * ag1[i] + (1.0 - r) * apg[i];
        double b = -1.0 / 3 * ah1[i] + (1.0 - r) * aph[i];
        macro_apply(p.xi + 3 * k, p.wv[k], U, al, be, ga, de, a, b);
        ag0[i] = a;
        ah0[i] = b;
    });
}

}  // namespace adj
