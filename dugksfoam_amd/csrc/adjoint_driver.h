// adjoint_driver.h -- mesh preparation, memory plan and the checkpointed forward/reverse sweep of
// the discrete adjoint, written once over a backend B:
//   B::alloc<T>(n)            n zeroed elements of device memory (nullptr and B::fail on error)
//   B::upload(d, h, bytes)    host -> device;  B::download(h, d, bytes) device -> host (waits)
//   B::copy(d, s, bytes)      device -> device;  B::zero(d, bytes)
//   B::launch<Tag>(n, fn)     fn(x, i) for points i < n, x the backend's executor (adjoint_core.h);
//                             Tag names the phase
//   B::free_bytes()           device memory available for the trajectory
//   B::tic(slot) / B::ms(a, b) device-time marks
//   B::sync(), B::bad, B::msg error state; B::clear_error() forgets a failed optional allocation
// dugks_adjoint.hip instantiates it with the HIP backend; the CPU check build instantiates it with
// plain loops.
//
// Memory plan: the forward sweep keeps a checkpoint of the state every C steps; the reverse sweep
// walks the segments backwards, re-running each segment's forward steps into a tape of C - 1 states
// and reversing them step by step.  States hold g~, h~, the lagged rows and the macros, so
// (ceil(N / C) + C + 1) states are resident at once (C = ceil(sqrt N) unless given).
#pragma once

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "adjoint_core.h"
#include "lsq.h"

#ifdef __HIPCC__
#define ADJ_FN [=] __host__ __device__(auto &x, int i)
#else
#define ADJ_FN [=](auto &x, int i)
#endif

namespace adj {

// Launch tags: they name the kernel of each phase in profiles.
struct k_wall_in {};
struct k_init {};
struct k_s1 {};
struct k_grad {};
struct k_face {};
struct k_cell {};
struct k_face_adj {};
struct k_recon_adj {};
struct k_grad_adj {};
struct k_s1_adj {};

struct Mesh {
    int D = 0, K = 0, nc = 0, nfi = 0, nf = 0, nv = 0;
    double R = 0, omega = 0, Tref = 0, mu = 0, Pr = 0;
    std::vector<int> cstart, cface, owner, nei, fkind, partner, pcell;
    std::vector<double> S, Cf, C, V, wf, pvec, nvec, nhat, adel, xi, wv, Tw, Uw, rb, pb, muc;
    std::vector<int> sax, mir;
};

// Validates the case and derives the per-face tables.  Returns nullptr or the reason (code set).
inline const char *prepare(const dugks_case *cs, Mesh &m, int &code)
{
    code = DUGKS_ERR_INVALID;
    if (!cs) return "null case";
    if (cs->abi_version != DUGKS_ABI_VERSION) return "abi_version mismatch";
    if (cs->dim < 1 || cs->dim > 3) return "dim must be 1, 2 or 3";
    if (cs->n_cells <= 0 || cs->n_internal_faces < 0 || cs->n_faces < cs->n_internal_faces) return "bad mesh sizes";
    if (!cs->owner || (cs->n_internal_faces && !cs->neighbour) || !cs->C || !cs->V || !cs->Sf || !cs->Cf ||
        !cs->weights || (cs->n_faces > cs->n_internal_faces && !cs->delta))
        return "missing mesh array";
    if (cs->n_xi <= 0 || cs->n_xi > 255 || !cs->xi || !cs->xi_weights) return "bad velocity set";
    if (!(cs->R > 0) || !(cs->T_ref > 0) || !(cs->mu_ref > 0) || cs->K < 0) return "bad gas properties";
    const int nc = cs->n_cells, nfi = cs->n_internal_faces, nf = cs->n_faces, nb = nf - nfi;
    // batches: one muRef per case (the time step is shared, as in dugks_step)
    m.muc.assign(nc, cs->mu_ref);
    if (cs->n_cases > 1) {
        if (!cs->cell_case) return "n_cases > 1 without cell_case";
        for (int c = 0; c < nc; ++c) {
            const int k = cs->cell_case[c];
            if (k < 0 || k >= cs->n_cases) return "cell_case out of range";
            if (cs->mu_ref_case) m.muc[c] = cs->mu_ref_case[k];
            if (!(m.muc[c] > 0)) return "muRef must be > 0";
        }
    }
    code = DUGKS_ERR_UNSUPPORTED;
    m.D = cs->dim; m.K = cs->K; m.nc = nc; m.nfi = nfi; m.nf = nf;
    m.R = cs->R; m.omega = cs->omega; m.Tref = cs->T_ref; m.mu = cs->mu_ref; m.Pr = cs->Pr;
    const int n = cs->n_xi;
    m.nv = m.D == 3 ? n * n * n : (m.D == 2 ? n * n : n);
    std::vector<int> bkind(nb, 0), bpartner(nb, -1);
    m.Tw.assign(nf, 0.0);
    m.Uw.assign(3 * (size_t)nf, 0.0);
    m.rb.assign(nf, 0.0);
    m.sax.assign(nf, 0);
    m.pb.assign(nf, 0.0);
    for (int pi = 0; pi < cs->n_patches; ++pi) {
        const dugks_patch &pp = cs->patches[pi];
        code = DUGKS_ERR_INVALID;
        if (pp.start < nfi || pp.size < 0 || pp.start + pp.size > nf) return "patch face range outside the boundary";
        code = DUGKS_ERR_UNSUPPORTED;
        if (pp.type != DUGKS_PATCH_WALL && pp.type != DUGKS_PATCH_CYCLIC && pp.type != DUGKS_PATCH_FIXED &&
            pp.type != DUGKS_PATCH_ZEROGRADIENT && pp.type != DUGKS_PATCH_SYMMETRY &&
            pp.type != DUGKS_PATCH_SYMMETRY_PLANE &&
            pp.type != DUGKS_PATCH_FARFIELD && pp.type != DUGKS_PATCH_PRESSURE_IN &&
            pp.type != DUGKS_PATCH_PRESSURE_OUT)
            return "unknown patch type";
        int axis = 0;
        if ((pp.type == DUGKS_PATCH_SYMMETRY || pp.type == DUGKS_PATCH_SYMMETRY_PLANE) && pp.size) {
            // the mirror axis from the patch's first face normal (discreteVelocity.C:766-768)
            const double *S0 = cs->Sf + 3 * (size_t)pp.start;
            const double mS = dugks_lsq::mag3(S0);
            for (int d = 1; d < 3; ++d)
                if (std::fabs(S0[d]) > std::fabs(S0[axis])) axis = d;
            if (!(mS > 0) || !(std::fabs(S0[axis]) / mS > 1.0 - 1e-9) || axis >= cs->dim)
                return "symmetry patch must be normal to a solution axis";
        }
        code = DUGKS_ERR_INVALID;
        if (pp.type == DUGKS_PATCH_WALL && (!pp.U || !pp.T)) return "wall patch without U/T";
        if (pp.type == DUGKS_PATCH_FIXED && (!pp.U || !pp.T || !pp.rho)) return "fixed patch without rho, U, T";
        if (pp.type == DUGKS_PATCH_FARFIELD && (!pp.T || !pp.rho)) return "far-field patch without rho, T";
        if ((pp.type == DUGKS_PATCH_PRESSURE_IN || pp.type == DUGKS_PATCH_PRESSURE_OUT) &&
            (!pp.T || !(pp.pressure > 0)))
            return "pressure patch without T or a positive pressure";
        if (pp.type == DUGKS_PATCH_CYCLIC &&
            (pp.neighbour_patch < 0 || pp.neighbour_patch >= cs->n_patches ||
             cs->patches[pp.neighbour_patch].size != pp.size ||
             cs->patches[pp.neighbour_patch].type != DUGKS_PATCH_CYCLIC))
            return "cyclic patch without a matching partner";
        for (int i = 0; i < pp.size; ++i) {
            const int b = pp.start - nfi + i, f = pp.start + i;
            if (bkind[b]) return "boundary face in two patches";
            bkind[b] = pp.type;
            if (pp.type == DUGKS_PATCH_WALL || pp.type == DUGKS_PATCH_FIXED) {
                m.Tw[f] = pp.T[i];
                if (!(m.Tw[f] > 0)) return "boundary temperature must be > 0";
                for (int d = 0; d < 3; ++d) m.Uw[3 * (size_t)f + d] = pp.U[3 * i + d];
                if (pp.type == DUGKS_PATCH_FIXED) {
                    m.rb[f] = pp.rho[i];
                    if (!(m.rb[f] > 0)) return "boundary density must be > 0";
                }
            } else if (pp.type == DUGKS_PATCH_PRESSURE_IN || pp.type == DUGKS_PATCH_PRESSURE_OUT) {
                m.Tw[f] = pp.T[i];
                m.pb[f] = pp.pressure;
                if (!(m.Tw[f] > 0)) return "boundary temperature must be > 0";
            } else if (pp.type == DUGKS_PATCH_FARFIELD) {
                m.Tw[f] = pp.T[i];
                m.rb[f] = pp.rho[i];
                if (!(m.Tw[f] > 0) || !(m.rb[f] > 0)) return "boundary density and temperature must be > 0";
            } else if (pp.type == DUGKS_PATCH_ZEROGRADIENT) {
                // nothing beyond the lagged gradient row
            } else if (pp.type == DUGKS_PATCH_SYMMETRY || pp.type == DUGKS_PATCH_SYMMETRY_PLANE) {
                m.sax[f] = axis;
            } else {
                bpartner[b] = cs->patches[pp.neighbour_patch].start - nfi + i;
            }
        }
    }
    code = DUGKS_ERR_INVALID;
    for (int b = 0; b < nb; ++b)
        if (!bkind[b]) return "boundary face not in any patch";
    for (int f = 0; f < nf; ++f)
        if (cs->owner[f] < 0 || cs->owner[f] >= nc) return "owner out of range";
    for (int f = 0; f < nfi; ++f)
        if (cs->neighbour[f] < 0 || cs->neighbour[f] >= nc || cs->neighbour[f] == cs->owner[f])
            return "neighbour out of range";
    for (int c = 0; c < nc; ++c)
        if (!(cs->V[c] > 0) || !std::isfinite(cs->V[c])) return "non-positive cell volume";
    for (int f = 0; f < nf; ++f)
        if (!(dugks_lsq::mag3(cs->Sf + 3 * f) > 0)) return "zero-area face";
    if (const char *why = dugks_lsq::vectors(cs, bkind, m.pvec, m.nvec)) return why;

    m.owner.assign(cs->owner, cs->owner + nf);
    m.nei.assign(nfi ? cs->neighbour : nullptr, nfi ? cs->neighbour + nfi : nullptr);
    if (m.nei.empty()) m.nei.push_back(-1);
    m.S.assign(cs->Sf, cs->Sf + 3 * (size_t)nf);
    m.Cf.assign(cs->Cf, cs->Cf + 3 * (size_t)nf);
    m.C.assign(cs->C, cs->C + 3 * (size_t)nc);
    m.V.assign(cs->V, cs->V + nc);
    m.wf.assign(cs->weights, cs->weights + nf);
    if (m.nvec.empty()) m.nvec.assign(3, 0.0);
    m.fkind.assign(nf, FK_INTERNAL);
    m.partner.assign(nf, -1);
    m.pcell.assign(nf, -1);
    m.nhat.assign(3 * (size_t)nf, 0.0);
    m.adel.assign(nf, 0.0);
    for (int b = 0; b < nb; ++b) {
        const int f = nfi + b;
        if (bkind[b] != DUGKS_PATCH_CYCLIC) {
            m.fkind[f] = bkind[b] == DUGKS_PATCH_WALL    ? FK_WALL
                         : bkind[b] == DUGKS_PATCH_FIXED   ? FK_FIXED
                         : bkind[b] == DUGKS_PATCH_SYMMETRY ? FK_SYM
                         : bkind[b] == DUGKS_PATCH_SYMMETRY_PLANE ? FK_SPL
                         : bkind[b] == DUGKS_PATCH_FARFIELD ? FK_FAR
                         : bkind[b] == DUGKS_PATCH_PRESSURE_IN ? FK_PIN
                         : bkind[b] == DUGKS_PATCH_PRESSURE_OUT ? FK_POUT
                                                            : FK_ZG;
            const double mS = dugks_lsq::mag3(cs->Sf + 3 * f);
            for (int d = 0; d < 3; ++d) m.nhat[3 * (size_t)f + d] = cs->Sf[3 * f + d] / mS;
            m.adel[f] = 1.0 / (1.0 / dugks_lsq::mag3(cs->delta + 3 * b));
        } else {
            m.fkind[f] = FK_CYCLIC;
            m.partner[f] = nfi + bpartner[b];
            m.pcell[f] = cs->owner[nfi + bpartner[b]];
        }
    }
    // faces of each cell, ascending
    m.cstart.assign(nc + 1, 0);
    for (int f = 0; f < nf; ++f) m.cstart[cs->owner[f] + 1]++;
    for (int f = 0; f < nfi; ++f) m.cstart[cs->neighbour[f] + 1]++;
    for (int c = 0; c < nc; ++c) m.cstart[c + 1] += m.cstart[c];
    m.cface.assign(m.cstart[nc], 0);
    std::vector<int> fill(nc, 0);
    for (int f = 0; f < nf; ++f) {
        const int o = cs->owner[f];
        m.cface[m.cstart[o] + fill[o]++] = f;
        if (f < nfi) {
            const int nn = cs->neighbour[f];
            m.cface[m.cstart[nn] + fill[nn]++] = f;
        }
    }
    // velocity set in reference order (fvDVM.C:140-221)
    m.xi.assign(3 * (size_t)m.nv, 0.0);
    m.wv.assign(m.nv, 0.0);
    const double *X = cs->xi, *W = cs->xi_weights;
    int k = 0;
    const int ny = m.D >= 2 ? n : 1, nz = m.D == 3 ? n : 1;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ixx = 0; ixx < n; ++ixx, ++k) {
                m.wv[k] = m.D == 3 ? W[iz] * W[iy] * W[ixx] : (m.D == 2 ? W[iy] * W[ixx] * 1 : W[ixx] * 1 * 1);
                m.xi[3 * k] = X[ixx];
                m.xi[3 * k + 1] = m.D >= 2 ? X[iy] : 0.0;
                m.xi[3 * k + 2] = m.D == 3 ? X[iz] : 0.0;
            }
    // mirrored velocities (symmXtgID / symmYtgID / symmZtgID, fvDVM.C:152-166, 183-195): the 1-D
    // set is symmetric, so index i mirrors to n - 1 - i
    m.mir.assign(3 * (size_t)m.nv, 0);
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ixx = 0; ixx < n; ++ixx) {
                const int kk = (iz * ny + iy) * n + ixx;
                m.mir[kk] = (iz * ny + iy) * n + (n - 1 - ixx);
                m.mir[m.nv + kk] = (iz * ny + (ny - 1 - iy)) * n + ixx;
                m.mir[2 * m.nv + kk] = ((nz - 1 - iz) * ny + iy) * n + ixx;
            }
    code = DUGKS_OK;
    return nullptr;
}

template <class B>
struct Engine {
    B be;
    Mesh mesh;
    AP p{};
    int nb = 0;
    size_t ncv = 0, nfv = 0, nbv = 0;
    double *gp = nullptr, *hp = nullptr, *gg = nullptr, *gh = nullptr, *gf = nullptr, *hf = nullptr;
    double *abf = nullptr, *abh = nullptr;   // adjoint face values; relaxed face values in a forward step
    double *ag[2] = {nullptr, nullptr}, *ah[2] = {nullptr, nullptr};
    double *AGg[2] = {nullptr, nullptr}, *AGh[2] = {nullptr, nullptr};
    double *apg = nullptr, *aph = nullptr, *agg = nullptr, *agh = nullptr;
    FaceRec *fr = nullptr;
    double *win = nullptr, *dwin = nullptr, *jacc = nullptr, *dTw = nullptr, *coef = nullptr;
    double *aUf = nullptr;   // far-field faces: adjoint of the owner cell's U, rho, T [nf][5]
    double *rho0 = nullptr, *U0 = nullptr, *T0 = nullptr, *finmac = nullptr;
    double *states = nullptr;
    size_t states_n = 0;
    double *ftape = nullptr;   // face values and records of a segment's re-run steps
    size_t ftape_n = 0;
    int ftape_slots = 0;
    bool have_run = false;
    double fwd_ms = 0.0, rev_ms = 0.0;
    std::string err;

    int fail(int code, const std::string &m)
    {
        err = m;
        return code;
    }

    template <class T>
    T *up(const std::vector<T> &v)
    {
        T *d = be.template alloc<T>(v.size());
        if (d) be.upload(d, v.data(), v.size() * sizeof(T));
        return d;
    }

    int create(const dugks_case *cs)
    {
        int code;
        if (const char *why = prepare(cs, mesh, code)) return fail(code, why);
        const Mesh &m = mesh;
        nb = m.nf - m.nfi;
        ncv = (size_t)m.nc * m.nv;
        nfv = (size_t)m.nf * m.nv;
        nbv = (size_t)std::max(nb, 1) * m.nv;
        p.D = m.D; p.K = m.K; p.nc = m.nc; p.nfi = m.nfi; p.nf = m.nf; p.nv = m.nv;
        p.R = m.R; p.omega = m.omega; p.Tref = m.Tref; p.mu = m.mu; p.Pr = m.Pr;
        p.kc = m.K + 3.0 - m.D;
        p.cstart = up(m.cstart); p.cface = up(m.cface); p.owner = up(m.owner); p.nei = up(m.nei);
        p.fkind = up(m.fkind); p.partner = up(m.partner); p.pcell = up(m.pcell);
        p.S = up(m.S); p.Cf = up(m.Cf); p.C = up(m.C); p.V = up(m.V); p.wf = up(m.wf);
        p.pvec = up(m.pvec); p.nvec = up(m.nvec); p.nhat = up(m.nhat); p.adel = up(m.adel);
        p.xi = up(m.xi); p.wv = up(m.wv); p.Tw = up(m.Tw); p.Uw = up(m.Uw); p.rb = up(m.rb);
        p.sax = up(m.sax); p.mir = up(m.mir); p.pb = up(m.pb); p.muc = up(m.muc);
        win = be.template alloc<double>(m.nf);
        dwin = be.template alloc<double>(m.nf);
        p.win = win;
        p.dwin = dwin;
        const size_t ngv = (size_t)m.nc * m.D * m.nv;
        gp = be.template alloc<double>(ncv); hp = be.template alloc<double>(ncv);
        gg = be.template alloc<double>(ngv); gh = be.template alloc<double>(ngv);
        gf = be.template alloc<double>(nfv); hf = be.template alloc<double>(nfv);
        abf = be.template alloc<double>(nfv); abh = be.template alloc<double>(nfv);
        for (int i = 0; i < 2; ++i) {
            ag[i] = be.template alloc<double>(ncv); ah[i] = be.template alloc<double>(ncv);
            AGg[i] = be.template alloc<double>(nbv); AGh[i] = be.template alloc<double>(nbv);
        }
        apg = be.template alloc<double>(ncv); aph = be.template alloc<double>(ncv);
        agg = be.template alloc<double>(ngv); agh = be.template alloc<double>(ngv);
        fr = be.template alloc<FaceRec>(m.nf);
        jacc = be.template alloc<double>(m.nf); dTw = be.template alloc<double>(m.nf);
        aUf = be.template alloc<double>(5 * (size_t)m.nf);
        coef = be.template alloc<double>(m.nf);
        rho0 = be.template alloc<double>(m.nc); U0 = be.template alloc<double>(3 * (size_t)m.nc);
        T0 = be.template alloc<double>(m.nc); finmac = be.template alloc<double>(8 * (size_t)m.nc);
        be.sync();
        if (be.bad) return fail(DUGKS_ERR_MEMORY, "device allocation failed: " + be.msg);
        return DUGKS_OK;
    }

    size_t state_doubles() const { return 2 * ncv + 2 * nbv + 8 * (size_t)mesh.nc; }

    St state(size_t i) const
    {
        double *b = states + i * state_doubles();
        St s;
        s.g = b;
        s.h = b + ncv;
        s.Gg = b + 2 * ncv;
        s.Gh = b + 2 * ncv + nbv;
        s.mac = b + 2 * ncv + 2 * nbv;
        return s;
    }

    void copy_state(St d, St s) { be.copy(d.g, s.g, state_doubles() * sizeof(double)); }

    size_t fslot_doubles() const { return 2 * nfv + 32 * (size_t)mesh.nf; }
    double *fslot_g(int s) const { return ftape + (size_t)s * fslot_doubles(); }
    double *fslot_h(int s) const { return fslot_g(s) + nfv; }
    FaceRec *fslot_r(int s) const { return reinterpret_cast<FaceRec *>(fslot_g(s) + 2 * nfv); }

    // one forward step's phases from state a; with fwd, the new state into b.  The face values
    // and records go to the work arrays, or to face-tape slot `slot` when it is >= 0.
    void recompute(St a, St b, int step, bool fwd, bool j, int slot = -1)
    {
        const AP q = p;
        double *gp_ = gp, *hp_ = hp, *gg_ = gg, *gh_ = gh;
        double *gf_ = slot >= 0 ? fslot_g(slot) : gf, *hf_ = slot >= 0 ? fslot_h(slot) : hf;
        FaceRec *fr_ = slot >= 0 ? fslot_r(slot) : fr;
        double *gr = fwd ? abf : nullptr, *hr = fwd ? abh : nullptr;
        const double *cf = coef;
        double *ja = j ? jacc : nullptr;
        double *Gng = fwd ? b.Gg : nullptr, *Gnh = fwd ? b.Gh : nullptr;
        be.template launch<k_s1>(q.nc, ADJ_FN { s1_cell(x, q, a, i, gp_, hp_); });
        be.template launch<k_grad>(q.nc, ADJ_FN { grad_cell(x, q, a, i, gp_, hp_, gg_, gh_, Gng, Gnh); });
        const double *mac = a.mac;
        be.template launch<k_face>(q.nf, ADJ_FN { face_fwd(x, q, i, mac, gp_, hp_, gg_, gh_, gf_, hf_, fr_, gr, hr, cf, ja); });
        if (fwd) be.template launch<k_cell>(q.nc, ADJ_FN { cell_fwd(x, q, a, i, gp_, hp_, gr, hr, b); });
    }

    // reverse of step n (state a = state n); cur: index of the adjoint of state n + 1.  The
    // step's face values and records come from face-tape slot `slot` (>= 0) or are recomputed.
    void reverse(St a, int n, bool seed, int &cur, int slot = -1)
    {
        if (slot < 0) recompute(a, a, n, false, false);
        const AP q = p;
        const FaceRec *fr_ = slot >= 0 ? fslot_r(slot) : fr;
        const double *gf_ = slot >= 0 ? fslot_g(slot) : gf, *hf_ = slot >= 0 ? fslot_h(slot) : hf;
        const double *cf = seed ? coef : nullptr;
        const double *ag1 = ag[cur], *ah1 = ah[cur], *AG1g = AGg[cur], *AG1h = AGh[cur];
        double *ag0 = ag[cur ^ 1], *ah0 = ah[cur ^ 1], *AG0g = AGg[cur ^ 1], *AG0h = AGh[cur ^ 1];
        double *abf_ = abf, *abh_ = abh, *dTw_ = dTw, *apg_ = apg, *aph_ = aph, *agg_ = agg, *agh_ = agh;
        double *aUf_ = aUf;
        const double *mac = a.mac;
        be.template launch<k_face_adj>(q.nf, ADJ_FN { face_adj(x, q, i, mac, fr_, gf_, hf_, ag1, ah1, cf, abf_, abh_, dTw_, aUf_); });
        if (n == 0) return;   // the initial state's adjoint is not needed
        be.template launch<k_recon_adj>(q.nc, ADJ_FN { recon_adj(x, q, i, abf_, abh_, ag1, ah1, AG1g, AG1h, apg_, aph_, agg_, agh_); });
        be.template launch<k_grad_adj>(q.nc, ADJ_FN { grad_adj(x, q, i, agg_, agh_, apg_, aph_, AG0g, AG0h); });
        be.template launch<k_s1_adj>(q.nc, ADJ_FN { s1_adj(x, q, a, i, ag1, ah1, apg_, aph_, aUf_, ag0, ah0); });
        cur ^= 1;
    }

    int run(const double *rho, const double *U, const double *T, double dt, int nsteps, int first, int every,
            const double *face_coef, double *J, double *dJdT)
    {
        const Mesh &m = mesh;
        if (!rho || !U || !T || !face_coef) return fail(DUGKS_ERR_INVALID, "null input");
        if (!(dt > 0) || !std::isfinite(dt)) return fail(DUGKS_ERR_INVALID, "dt must be > 0");
        if (nsteps < 1 || every < 0) return fail(DUGKS_ERR_INVALID, "n_steps must be >= 1, checkpoint_every >= 0");
        for (int c = 0; c < m.nc; ++c)
            if (!(rho[c] > 0) || !(T[c] > 0)) return fail(DUGKS_ERR_INVALID, "initial rho and T must be > 0");
        const int C = std::min(nsteps, every > 0 ? every : (int)std::ceil(std::sqrt((double)nsteps)));
        const int ncp = (nsteps + C - 1) / C;
        const size_t nst = (size_t)ncp + (size_t)C + 1;   // checkpoints, tape (C - 1), cur, nxt
        const size_t need = nst * state_doubles();
        if (need != states_n) {
            if (states) be.release(states);
            states = nullptr;
            states_n = 0;
            if ((double)need * sizeof(double) > 0.9 * be.free_bytes())
                return fail(DUGKS_ERR_MEMORY, "trajectory of " + std::to_string(nst) + " states (" +
                                                  std::to_string((double)need * 8.0 / 1e9) +
                                                  " GB) does not fit; raise checkpoint spacing or shorten the run");
            states = be.template alloc<double>(need);
            if (!states) return fail(DUGKS_ERR_MEMORY, "device allocation failed: " + be.msg);
            states_n = need;
        }
        // face tape: the face values and records of a segment's re-run steps, so that their
        // reverse steps skip the recomputation; taken only where it fits beside the trajectory
        {
            const size_t want = (size_t)(C - 1) * fslot_doubles();
            if (want != ftape_n) {
                if (ftape) be.release(ftape);
                ftape = nullptr;
                ftape_n = 0;
                ftape_slots = 0;
                if (want && (double)want * sizeof(double) < 0.8 * be.free_bytes()) {
                    ftape = be.template alloc<double>(want);
                    if (ftape) {
                        ftape_n = want;
                        ftape_slots = C - 1;
                    } else {
                        be.clear_error();   // an optional buffer: recompute instead
                    }
                }
            }
        }
        p.dt = dt;
        be.upload(rho0, rho, m.nc * sizeof(double));
        be.upload(U0, U, 3 * (size_t)m.nc * sizeof(double));
        be.upload(T0, T, m.nc * sizeof(double));
        be.upload(coef, face_coef, m.nf * sizeof(double));
        be.zero(jacc, m.nf * sizeof(double));
        be.zero(dTw, m.nf * sizeof(double));
        const AP q = p;
        double *win_ = win, *dwin_ = dwin;
        const double *r0 = rho0, *u0 = U0, *t0 = T0;
        const int nfi = m.nfi;
        be.template launch<k_wall_in>(nb, ADJ_FN {
            if (q.fkind[nfi + i] == FK_WALL) wall_in(x, q, nfi + i, win_, dwin_);
        });
        St cur = state(ncp + C - 1), nxt = state(ncp + C);
        be.template launch<k_init>(q.nc, ADJ_FN { init_cell(x, q, cur, i, r0, u0, t0); });
        be.zero(cur.Gg, 2 * nbv * sizeof(double));
        be.tic(0);
        for (int n = 0; n < nsteps; ++n) {
            if (n % C == 0) copy_state(state(n / C), cur);
            recompute(cur, nxt, n, true, n + 1 >= first);
            std::swap(cur, nxt);
        }
        be.tic(1);
        be.copy(finmac, cur.mac, 8 * (size_t)m.nc * sizeof(double));
        std::vector<double> jh(m.nf);
        be.download(jh.data(), jacc, m.nf * sizeof(double));
        double Jsum = 0.0;
        for (int f = 0; f < m.nf; ++f) Jsum += jh[f];
        int ca = 0;
        be.zero(ag[0], ncv * sizeof(double));
        be.zero(ah[0], ncv * sizeof(double));
        be.zero(AGg[0], nbv * sizeof(double));
        be.zero(AGh[0], nbv * sizeof(double));
        be.tic(2);
        for (int seg = ncp - 1; seg >= 0; --seg) {
            const int n0 = seg * C, n1 = std::min(nsteps, n0 + C);
            auto tape = [&](int n) { return n == n0 ? state(seg) : state(ncp + (n - n0 - 1)); };
            auto slot = [&](int n) { return n + 1 < n1 && n - n0 < ftape_slots ? n - n0 : -1; };
            for (int n = n0; n + 1 < n1; ++n) recompute(tape(n), tape(n + 1), n, true, false, slot(n));
            for (int n = n1 - 1; n >= n0; --n) reverse(tape(n), n, n + 1 >= first, ca, slot(n));
        }
        be.tic(3);
        std::vector<double> dh(m.nf);
        be.download(dh.data(), dTw, m.nf * sizeof(double));
        be.sync();
        if (be.bad) return fail(DUGKS_ERR_DEVICE, "device error: " + be.msg);
        fwd_ms = be.ms(0, 1);
        rev_ms = be.ms(2, 3);
        if (!std::isfinite(Jsum)) return fail(DUGKS_ERR_NONFINITE, "objective is not finite");
        if (J) *J = Jsum;
        if (dJdT)
            for (int b = 0; b < nb; ++b) dJdT[b] = dh[m.nfi + b];
        have_run = true;
        return DUGKS_OK;
    }

    int macro(double *rho, double *U, double *T, double *q)
    {
        if (!have_run) return fail(DUGKS_ERR_STATE, "no run yet");
        const int nc = mesh.nc;
        std::vector<double> mh(8 * (size_t)nc);
        be.download(mh.data(), finmac, mh.size() * sizeof(double));
        if (be.bad) return fail(DUGKS_ERR_DEVICE, "device error: " + be.msg);
        for (int c = 0; c < nc; ++c) {
            const double *a = &mh[8 * (size_t)c];
            if (rho) rho[c] = a[0];
            if (T) T[c] = a[4];
            for (int d = 0; d < 3; ++d) {
                if (U) U[3 * c + d] = a[1 + d];
                if (q) q[3 * c + d] = a[5 + d];
            }
        }
        return DUGKS_OK;
    }
};

}  // namespace adj
