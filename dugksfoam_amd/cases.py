"""Case front-end: meshes, geometry, discrete-velocity sets and gas data.

Builds the inputs that dugksFoam reads from an OpenFOAM case directory
(polyMesh + fvMesh geometry, constant/DVMProperties, constant/Xis,
constant/weights, 0/{rho,U,T}) as plain numpy arrays, in OpenFOAM's mesh
conventions, so that the engine and the oracle see the same data.

Geometry follows OpenFOAM's definitions:
  * internal faces in upper-triangular order (owner ascending, then
    neighbour), Sf pointing owner -> neighbour; boundary faces patch by patch
    with Sf outward; empty patches omitted (fvPatch size 0);
  * weights w_f = |Sf.(C_N - Cf)| / (|Sf.(Cf - C_P)| + |Sf.(C_N - Cf)|)
    (surfaceInterpolation::makeWeights); 1 on non-coupled patches;
  * boundary delta = n (n . (Cf - C_P)) on non-coupled patches
    (fvPatch::delta), (Cf - C_P) - (Cf' - C_N) on cyclic patches.

Velocity sets restate src/scripts/setDV.py:83-138 and the gas parameters
demo/cavity/para.py:33-46.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

PATCH_WALL = 1
PATCH_CYCLIC = 2
PATCH_SYMMETRY = 3
PATCH_FARFIELD = 4
PATCH_FIXED = 5
PATCH_PRESSURE_IN = 6
PATCH_PRESSURE_OUT = 7
PATCH_ZEROGRADIENT = 8
PATCH_SYMMETRY_PLANE = 9

# Boltzmann constant and argon data of demo/cavity/para.py:13-20
K_BOLTZMANN = 1.38066e-23
M_ARGON = 6.63e-26
D_ARGON = 4.17e-10


# ---------------------------------------------------------------------------
# discrete velocity sets
# ---------------------------------------------------------------------------

def dv_newton_cotes(xi_max: float, n: int):
    """1-D uniform velocity grid on [-xi_max, xi_max] with n points.

    n = 4Z+1: compound Newton-Cotes (Boole) weights exactly as
    src/scripts/setDV.py:114-138 (dvNC): interior pattern 14,32,12,32 /90*4,
    end points 7/90*4, all times the spacing.
    Other n (the rule setDV.py leaves undefined): composite trapezoid on the
    same uniform nodes, the closed Newton-Cotes rule of degree 1, which keeps
    the weights mirror-symmetric.
    """
    n = int(n)
    if n < 3:
        raise ValueError("need at least 3 velocity points per axis")
    xi_min = -xi_max
    dv = (xi_max - xi_min) / (n - 1)
    xis = np.array([xi_min + dv * i for i in range(n)], dtype=np.float64)
    w = np.zeros(n)
    if n % 4 == 1:
        for i in range((n - 1) // 4):
            w[4 * i + 0] = 14.0 / 90 * 4
            w[4 * i + 1] = 32.0 / 90 * 4
            w[4 * i + 2] = 12.0 / 90 * 4
            w[4 * i + 3] = 32.0 / 90 * 4
        w[0] = 7.0 / 90 * 4
        w[n - 1] = 7.0 / 90 * 4
        w = dv * w
    else:
        w[:] = dv
        w[0] = 0.5 * dv
        w[n - 1] = 0.5 * dv
    return xis, w


def dv_gauss_hermite_half(c: float, n2: int):
    """Half-range Gauss-Hermite velocity set of src/scripts/setDV.py:83-112 (dvGH).

    Golub-Welsch on the recurrence of the half-range Hermite weight exp(-x^2)
    on [0, inf), mirrored to the negative axis; weights rescaled by exp(x^2)
    and by the most probable speed c.
    """
    n2 = int(n2)
    if n2 % 2 or n2 < 6:
        raise ValueError("half-range Gauss-Hermite needs an even count >= 6")
    n = n2 // 2
    a = np.zeros(n)
    b = np.zeros(n)
    a[0] = 1.0 / np.sqrt(np.pi)
    a[1] = 2.0 / np.sqrt(np.pi) / (np.pi - 2.0)
    b[1] = a[0] / (a[0] + a[1]) / 2.0
    for i in range(2, n):
        b[i] = (i - 1) + 1.0 / 2.0 - b[i - 1] - a[i - 1] ** 2
        a[i] = (i ** 2 / 4.0 / b[i] - b[i - 1] - 1.0 / 2) / a[i - 1] - a[i - 1]
    jac = np.diag(a) + np.diag(np.sqrt(b[1:n]), 1) + np.diag(np.sqrt(b[1:n]), -1)
    v, vec = np.linalg.eigh(jac)
    w = vec[0, :] * vec[0, :] * np.sqrt(np.pi) / 2.0
    order = np.argsort(v)
    v, w = v[order], w[order]
    xis = np.hstack((-np.flipud(v), v))
    ws = np.hstack((np.flipud(w), w))
    ws = ws * np.exp(xis ** 2) * c
    return xis * c, ws


# ---------------------------------------------------------------------------
# gas
# ---------------------------------------------------------------------------

def mu_from_kn(kn: float, R: float, T: float, rho: float, L: float = 1.0,
               alpha: float = 1.0, omega: float = 0.5) -> float:
    """Reference viscosity for a Knudsen number (demo/cavity/para.py:33-46)."""
    mu_bar = (5.0 * (alpha + 1.0) * (alpha + 2.0) * math.sqrt(math.pi) * kn
              / (4.0 * alpha * (5.0 - 2 * omega) * (7.0 - 2.0 * omega)))
    c = math.sqrt(2 * R * T)
    return rho * c * L * mu_bar


def argon_density(kn: float, L: float = 1.0) -> float:
    """Hard-sphere argon density for a Knudsen number (para.py:36-40)."""
    mfp = L * kn
    return M_ARGON / (mfp * math.sqrt(2.0) * math.pi * D_ARGON * D_ARGON)


# ---------------------------------------------------------------------------
# mesh containers
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Patch:
    name: str
    type: int
    owner: np.ndarray          # [n] int
    Sf: np.ndarray             # [n,3]
    Cf: np.ndarray             # [n,3]
    U: Optional[np.ndarray] = None   # [n,3]
    T: Optional[np.ndarray] = None   # [n]
    neighbour_patch: int = -1
    rho: Optional[np.ndarray] = None  # [n] far-field / fixed patches
    pressure: float = 0.0             # pressureIn / pressureOut patches


@dataclasses.dataclass
class Case:
    """One case: mesh + geometry + velocity set + gas + initial fields."""
    name: str
    dim: int
    C: np.ndarray
    V: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    Sf: np.ndarray
    Cf: np.ndarray
    weights: np.ndarray
    delta: np.ndarray
    patches: List[dict]
    xi: np.ndarray
    xi_weights: np.ndarray
    xi_max: float
    R: float
    omega: float
    T_ref: float
    mu_ref: float
    Pr: float
    K: int
    rho0: np.ndarray
    U0: np.ndarray
    T0: np.ndarray
    dt: float
    n_cases: int = 1
    cell_case: Optional[np.ndarray] = None
    mu_ref_case: Optional[np.ndarray] = None
    dt_case: Optional[np.ndarray] = None
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return int(self.V.shape[0])

    @property
    def n_internal_faces(self) -> int:
        return int(self.neighbour.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def nv(self) -> int:
        return int(self.xi.shape[0]) ** self.dim

    def velocities(self):
        """Tensor-product velocity set in reference order (fvDVM.C:140-221)."""
        n = self.xi.shape[0]
        if self.dim == 3:
            iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
            ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
            xi = np.stack([self.xi[ix], self.xi[iy], self.xi[iz]], axis=1)
            w = self.xi_weights[iz] * self.xi_weights[iy] * self.xi_weights[ix]
        elif self.dim == 2:
            iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            ix, iy = ix.ravel(), iy.ravel()
            xi = np.stack([self.xi[ix], self.xi[iy], np.zeros(ix.size)], axis=1)
            w = self.xi_weights[iy] * self.xi_weights[ix] * 1
        else:
            ix = np.arange(n)
            xi = np.stack([self.xi[ix], np.zeros(n), np.zeros(n)], axis=1)
            w = self.xi_weights[ix] * 1 * 1
        return xi, w


def _assemble(name, dim, C, V, own_i, nei_i, Sf_i, Cf_i, patches: List[Patch]) -> dict:
    """Concatenate internal and boundary faces and derive weights and deltas."""
    own = [own_i]
    Sf = [Sf_i]
    Cf = [Cf_i]
    pl = []
    start = own_i.shape[0]
    for p in patches:
        n = p.owner.shape[0]
        own.append(p.owner)
        Sf.append(p.Sf)
        Cf.append(p.Cf)
        pl.append(dict(name=p.name, type=p.type, start=start, size=n,
                       neighbour_patch=p.neighbour_patch,
                       U=None if p.U is None else np.ascontiguousarray(p.U, dtype=np.float64),
                       T=None if p.T is None else np.ascontiguousarray(p.T, dtype=np.float64),
                       rho=None if p.rho is None else np.ascontiguousarray(p.rho, dtype=np.float64),
                       pressure=float(p.pressure)))
        start += n
    owner = np.concatenate(own).astype(np.int32)
    Sf = np.concatenate(Sf).astype(np.float64)
    Cf = np.concatenate(Cf).astype(np.float64)
    nf = own_i.shape[0]
    weights = np.ones(owner.shape[0])
    # internal weights (surfaceInterpolation::makeWeights)
    sfd_own = np.abs(np.einsum("ij,ij->i", Sf[:nf], Cf[:nf] - C[own_i]))
    sfd_nei = np.abs(np.einsum("ij,ij->i", Sf[:nf], C[nei_i] - Cf[:nf]))
    weights[:nf] = sfd_nei / (sfd_own + sfd_nei)
    nb = owner.shape[0] - nf
    delta = np.zeros((nb, 3))
    for p in pl:
        s, n = p["start"], p["size"]
        b0 = s - nf
        o = owner[s:s + n]
        if p["type"] == PATCH_CYCLIC:
            q = pl[p["neighbour_patch"]]
            qs = q["start"]
            d_own = Cf[s:s + n] - C[o]
            d_nbr = Cf[qs:qs + n] - C[owner[qs:qs + n]]
            delta[b0:b0 + n] = d_own - d_nbr
        else:
            nhat = Sf[s:s + n] / np.linalg.norm(Sf[s:s + n], axis=1)[:, None]
            delta[b0:b0 + n] = nhat * np.einsum("ij,ij->i", nhat, Cf[s:s + n] - C[o])[:, None]
    for p in pl:
        if p["type"] == PATCH_CYCLIC:
            # cyclicFvPatch::makeWeights: w = dni / (di + dni), d = nf . delta_coupled
            s, n = p["start"], p["size"]
            q = pl[p["neighbour_patch"]]
            qs = q["start"]
            nh = Sf[s:s + n] / np.linalg.norm(Sf[s:s + n], axis=1)[:, None]
            nq = Sf[qs:qs + n] / np.linalg.norm(Sf[qs:qs + n], axis=1)[:, None]
            di = np.einsum("ij,ij->i", nh, Cf[s:s + n] - C[owner[s:s + n]])
            dni = np.einsum("ij,ij->i", nq, Cf[qs:qs + n] - C[owner[qs:qs + n]])
            weights[s:s + n] = dni / (di + dni)
    return dict(name=name, dim=dim, C=np.ascontiguousarray(C, dtype=np.float64),
                V=np.ascontiguousarray(V, dtype=np.float64), owner=owner,
                neighbour=nei_i.astype(np.int32), Sf=Sf, Cf=Cf, weights=weights,
                delta=delta, patches=pl)


def box_mesh(nx: int, ny: int, nz: int = 0, Lx: float = 1.0, Ly: float = 1.0,
             Lz: float = 1.0, dz2d: float = 0.01):
    """Uniform box of nx*ny (2-D, one cell thick, nz=0) or nx*ny*nz hexes.

    Returns cell centres/volumes, the internal faces in OpenFOAM order and a
    dict of the boundary faces per side ('x0','x1','y0','y1','z0','z1'), each
    ordered along the side's cell numbering.
    """
    dim = 3 if nz > 0 else 2
    nzc = nz if nz > 0 else 1
    dx, dy = Lx / nx, Ly / ny
    dz = Lz / nz if nz > 0 else dz2d
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nzc), indexing="ij")
    # cell index c = i + nx*(j + ny*k): build arrays in that order
    c_of = lambda ii, jj, kk: ii + nx * (jj + ny * kk)
    ii = np.tile(np.arange(nx), ny * nzc)
    jj = np.tile(np.repeat(np.arange(ny), nx), nzc)
    kk = np.repeat(np.arange(nzc), nx * ny)
    C = np.stack([(ii + 0.5) * dx, (jj + 0.5) * dy, (kk + 0.5) * dz], axis=1)
    V = np.full(C.shape[0], dx * dy * dz)
    own, nei, Sf, Cf = [], [], [], []
    ncell = nx * ny * nzc
    cidx = np.arange(ncell)
    # per owner: +x, +y, +z faces in that order (neighbour ascending)
    fx = ii < nx - 1
    fy = jj < ny - 1
    fz = (kk < nzc - 1) if dim == 3 else np.zeros(ncell, bool)
    rows = []
    for c in range(ncell):
        if fx[c]:
            rows.append((c, c + 1, 0))
        if fy[c]:
            rows.append((c, c + nx, 1))
        if fz[c]:
            rows.append((c, c + nx * ny, 2))
    rows = np.array(rows, dtype=np.int64).reshape(-1, 3)
    own = rows[:, 0]
    nei = rows[:, 1]
    ax = rows[:, 2]
    area = np.array([dy * dz, dx * dz, dx * dy])
    Sf = np.zeros((rows.shape[0], 3))
    Sf[np.arange(rows.shape[0]), ax] = area[ax]
    Cf = 0.5 * (C[own] + C[nei])
    sides = {}

    def side(mask, axis, sign, coord):
        o = cidx[mask]
        s = np.zeros((o.size, 3))
        s[:, axis] = sign * area[axis]
        cf = C[o].copy()
        cf[:, axis] = coord
        return dict(owner=o, Sf=s, Cf=cf)

    sides["x0"] = side(ii == 0, 0, -1.0, 0.0)
    sides["x1"] = side(ii == nx - 1, 0, 1.0, Lx)
    sides["y0"] = side(jj == 0, 1, -1.0, 0.0)
    sides["y1"] = side(jj == ny - 1, 1, 1.0, Ly)
    if dim == 3:
        sides["z0"] = side(kk == 0, 2, -1.0, 0.0)
        sides["z1"] = side(kk == nzc - 1, 2, 1.0, Lz)
    return dim, C, V, own, nei, Sf, Cf, sides


def _wall(name, sd, U, T):
    n = sd["owner"].size
    Uw = np.tile(np.asarray(U, dtype=np.float64), (n, 1))
    Tw = np.full(n, float(T)) if np.isscalar(T) else np.asarray(T, dtype=np.float64)
    return Patch(name, PATCH_WALL, sd["owner"], sd["Sf"], sd["Cf"], Uw, Tw)


def _concat_sides(sds):
    return dict(owner=np.concatenate([s["owner"] for s in sds]),
                Sf=np.concatenate([s["Sf"] for s in sds]),
                Cf=np.concatenate([s["Cf"] for s in sds]))


def courant_dt(deltacoeff_max: float, dim: int, xi_max: float, co: float = 0.5) -> float:
    """Fixed time step for Courant number `co` with the gas at rest
    (fvDVM::getCoNum, fvDVM.C:1111-1119, with Usurf = 0)."""
    return co / (deltacoeff_max * math.sqrt(dim) * xi_max)


def _finish(geo: dict, xi, w, xi_max, R, omega, T_ref, mu_ref, Pr, K, rho0, T0, dt, **kw) -> Case:
    nc = geo["V"].shape[0]
    return Case(name=geo["name"], dim=geo["dim"], C=geo["C"], V=geo["V"], owner=geo["owner"],
                neighbour=geo["neighbour"], Sf=geo["Sf"], Cf=geo["Cf"], weights=geo["weights"],
                delta=geo["delta"], patches=geo["patches"], xi=np.asarray(xi, np.float64),
                xi_weights=np.asarray(w, np.float64), xi_max=float(xi_max), R=R, omega=omega,
                T_ref=T_ref, mu_ref=mu_ref, Pr=Pr, K=K,
                rho0=np.full(nc, float(rho0)), U0=np.zeros((nc, 3)), T0=np.full(nc, float(T0)),
                dt=dt, **kw)


GAMMA_MONATOMIC = 5.0 / 3.0


def cavity(n: int = 60, nv: int = 41, kn: float = 0.075, dim: int = 2, R: float = 208.0,
           T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 4.0, lid_mach: float = 0.16,
           Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5,
           T_wall: float = None, name: str = None) -> Case:
    """Lid-driven cavity on the unit square (dim=2, n*n quads) or cube (dim=3, n^3 hexes).

    Diffuse Maxwell walls at T0; the lid (y = 1) moves along +x at
    lid_mach*sqrt(gamma R T0).  Velocity grid: nv points per axis on
    [-xi_sigma, xi_sigma]*sqrt(R T0).  Viscosity from Kn with the hard-sphere
    formula of para.py (omega 0.5, L = 1).  Fixed dt at Courant number `co`.
    """
    if dim == 2:
        d, C, V, own, nei, Sf, Cf, sides = box_mesh(n, n)
        walls = [_wall("movingWall", sides["y1"], (0, 0, 0), T0),
                 _wall("fixedWalls", _concat_sides([sides["x0"], sides["x1"], sides["y0"]]), (0, 0, 0), T0)]
    else:
        d, C, V, own, nei, Sf, Cf, sides = box_mesh(n, n, n)
        walls = [_wall("movingWall", sides["y1"], (0, 0, 0), T0),
                 _wall("fixedWalls", _concat_sides([sides["x0"], sides["x1"], sides["y0"],
                                                     sides["z0"], sides["z1"]]), (0, 0, 0), T0)]
    u_lid = lid_mach * math.sqrt(GAMMA_MONATOMIC * R * T0)
    walls[0].U[:, 0] = u_lid
    if T_wall is not None:
        for wpatch in walls:
            wpatch.T[:] = T_wall
    geo = _assemble(name or f"cavity{dim}d_{n}_{nv}", d, C, V, own, nei, Sf, Cf, walls)
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(float(n), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, u_lid=u_lid, n=n, nv=nv))


def line_mesh(n: int, L: float = 1.0, dyz: float = 0.01):
    """1-D mesh: n cells on [0, L] along x, one cell thick in y and z (both empty directions,
    as a dugksFoam 1-D case has them).  Returns the box_mesh tuple with dim 1 and only the x
    sides."""
    d, C, V, own, nei, Sf, Cf, sides = box_mesh(n, 1, 0, L, dyz, dz2d=dyz)
    return 1, C, V, own, nei, Sf, Cf, {"x0": sides["x0"], "x1": sides["x1"]}


def fourier_flow(n: int = 100, nv: int = 40, kn: float = 0.1, T_left: float = 263.0, T_right: float = 283.0,
                 R: float = 208.0, T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 5.0,
                 Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5) -> Case:
    """1-D heat conduction between two parallel diffuse plates (Fourier flow): n cells on
    [0, 1], the plate at x = 0 at T_left and the plate at x = 1 at T_right, the gas at rest at T0
    to start.  A 1-D dugksFoam case (fvDVM.C:198-215: nv velocities along x)."""
    d, C, V, own, nei, Sf, Cf, sides = line_mesh(n)
    walls = [_wall("leftWall", sides["x0"], (0, 0, 0), T_left), _wall("rightWall", sides["x1"], (0, 0, 0), T_right)]
    geo = _assemble(f"fourier_{n}_{nv}", d, C, V, own, nei, Sf, Cf, walls)
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(float(n), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, n=n, nv=nv, T_left=T_left, T_right=T_right))


def shock_tube(n: int = 100, nv: int = 40, kn: float = 0.01, rho_ratio: float = 2.0, T_ratio: float = 1.25,
               R: float = 208.0, T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 6.0,
               Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5,
               ends: str = "zeroGradient") -> Case:
    """1-D shock tube on [0, 1]: gas at rest, rho0 and T0 * T_ratio for x < 0.5, rho0 / rho_ratio
    and T0 for x > 0.5; both ends zeroGradient patches (ends="zeroGradient") or diffuse walls at
    T0 (ends="wall").  The scheme reconstructs with unlimited least-squares gradients (the
    reference's limiter returns the unlimited gradient, VenkatakrishnanLimitedGrad.C:221-225), so
    a density ratio of 8 overshoots to a negative face temperature in the first step; 2 runs."""
    d, C, V, own, nei, Sf, Cf, sides = line_mesh(n)
    if ends == "zeroGradient":
        pl = [Patch("left", PATCH_ZEROGRADIENT, sides["x0"]["owner"], sides["x0"]["Sf"], sides["x0"]["Cf"]),
              Patch("right", PATCH_ZEROGRADIENT, sides["x1"]["owner"], sides["x1"]["Sf"], sides["x1"]["Cf"])]
    else:
        pl = [_wall("left", sides["x0"], (0, 0, 0), T0), _wall("right", sides["x1"], (0, 0, 0), T0)]
    geo = _assemble(f"shocktube_{n}_{nv}_{ends}", d, C, V, own, nei, Sf, Cf, pl)
    xi_max = xi_sigma * math.sqrt(R * T0 * T_ratio)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(float(n), d, xi_max, co)
    c = _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                meta=dict(kn=kn, n=n, nv=nv, ends=ends))
    left = c.C[:, 0] < 0.5
    c.rho0 = np.where(left, rho0, rho0 / rho_ratio)
    c.T0 = np.where(left, T0 * T_ratio, T0)
    return c


def agglomerated_cavity(n: int = 12, nv: int = 12, block: int = 3, kn: float = 0.075, R: float = 208.0,
                        T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 4.0, lid_mach: float = 0.16,
                        Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5) -> Case:
    """2-D lid-driven cavity on a general polygonal mesh: an n x n quad grid whose block x block
    groups of quads, on a checkerboard of blocks, are merged into single cells (polygons with
    4 block faces, several of them shared with one neighbour) beside plain quads.  Faces are the
    grid's edges between different cells, numbered in OpenFOAM's upper-triangular order
    (owner < neighbour, sorted by owner, then neighbour).  Walls and lid as in cavity()."""
    d, C4, V4, own4, nei4, Sf4, Cf4, sides = box_mesh(n, n)
    nb_ = (n + block - 1) // block
    qi, qj = np.arange(n * n) % n, np.arange(n * n) // n
    merged = ((qi // block + qj // block) % 2 == 0) & ((qi // block + 1) * block <= n) & ((qj // block + 1) * block <= n)
    key = np.where(merged, -1 - ((qj // block) * nb_ + qi // block), np.arange(n * n))
    cell_of = np.empty(n * n, dtype=np.int64)
    ids = {}
    for q in range(n * n):           # cells numbered in the order of their first quad
        k = int(key[q])
        if k not in ids:
            ids[k] = len(ids)
        cell_of[q] = ids[k]
    nc = len(ids)
    V = np.bincount(cell_of, weights=V4, minlength=nc)
    C = np.stack([np.bincount(cell_of, weights=V4 * C4[:, a], minlength=nc) for a in range(3)], axis=1) / V[:, None]
    a, b = cell_of[own4], cell_of[nei4]
    keep = a != b
    a, b, Sf, Cf = a[keep], b[keep], Sf4[keep].copy(), Cf4[keep]
    flip = a > b
    a, b = np.where(flip, b, a), np.where(flip, a, b)
    Sf[flip] *= -1.0
    order = np.lexsort((np.arange(a.size), b, a))
    own, nei, Sf, Cf = a[order], b[order], Sf[order], Cf[order]

    def remap(sd):
        return dict(owner=cell_of[sd["owner"]], Sf=sd["Sf"], Cf=sd["Cf"])

    walls = [_wall("movingWall", remap(sides["y1"]), (0, 0, 0), T0),
             _wall("fixedWalls", remap(_concat_sides([sides["x0"], sides["x1"], sides["y0"]])), (0, 0, 0), T0)]
    u_lid = lid_mach * math.sqrt(GAMMA_MONATOMIC * R * T0)
    walls[0].U[:, 0] = u_lid
    geo = _assemble(f"agglomerated_{n}_{block}_{nv}", d, C, V, own, nei, Sf, Cf, walls)
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(float(n), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, u_lid=u_lid, n=n, nv=nv, block=block))


def symmetry_cavity(nx: int = 30, ny: int = 60, nv: int = 28, kn: float = 0.075, R: float = 208.0,
                    T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 4.0, lid_mach: float = 0.16,
                    Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5,
                    plane: str = "x1", kind: str = "symmetryMod") -> Case:
    """Half of a 2-D cavity, nx x ny cells on [0, 0.5] x [0, 1], with a symmetry plane: the layout
    of the reference's demo/testSymmetry/symmetry case (lid y = 1 moving along +x, diffuse walls
    elsewhere, symmetryMod patch at x = 0.5).  plane = "x1" puts the plane on the right side
    (outward normal +x), "x0" on the left (normal -x).  kind = "symmetryPlane" makes the plane an
    OpenFOAM symmetryPlane patch (the serial form, doc/usage.tex:77-81) instead of symmetryMod."""
    d, C, V, own, nei, Sf, Cf, sides = box_mesh(nx, ny, 0, 0.5, 1.0)
    u_lid = lid_mach * math.sqrt(GAMMA_MONATOMIC * R * T0)
    other = "x0" if plane == "x1" else "x1"
    lid = _wall("movingWall", sides["y1"], (u_lid, 0, 0), T0)
    fixed = _wall("fixedWalls", _concat_sides([sides[other], sides["y0"]]), (0, 0, 0), T0)
    ptype = {"symmetryMod": PATCH_SYMMETRY, "symmetryPlane": PATCH_SYMMETRY_PLANE}[kind]
    sym = Patch("symmetryWall", ptype, sides[plane]["owner"], sides[plane]["Sf"], sides[plane]["Cf"])
    tag = "spl" if kind == "symmetryPlane" else ""
    geo = _assemble(f"symcavity{tag}_{nx}x{ny}_{nv}", d, C, V, own, nei, Sf, Cf, [lid, fixed, sym])
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(max(2.0 * nx, float(ny)), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, u_lid=u_lid, nx=nx, ny=ny, nv=nv))


def open_channel(nx: int = 20, ny: int = 8, Lx: float = 2.0, nv: int = 16, kn: float = 0.2,
                 rho_in: float = 1.05, u_in: float = 30.0, R: float = 208.0, T0: float = 273.0,
                 rho0: float = 1.0, xi_sigma: float = 4.0, Pr: float = 2.0 / 3.0, omega: float = 0.5,
                 K: int = 0, co: float = 0.5, outlet: str = "farField") -> Case:
    """Open 2-D channel [0, Lx] x [0, 1]: diffuse walls at y = 0, 1 (T0); the inlet x = 0 a fixed
    boundary Maxwellian (rho fixedValue: rho_in, U = (u_in, 0, 0), T0); the outlet x = Lx a far
    field (rho farField: rho0, T0, incoming velocity from the outlet cells) or, with
    outlet="zeroGradient", a zeroGradient patch (face values = the outlet cells' values)."""
    d, C, V, own, nei, Sf, Cf, sides = box_mesh(nx, ny, 0, Lx, 1.0)
    walls = _wall("walls", _concat_sides([sides["y0"], sides["y1"]]), (0, 0, 0), T0)
    n_in, n_out = sides["x0"]["owner"].size, sides["x1"]["owner"].size
    inlet = Patch("inlet", PATCH_FIXED, sides["x0"]["owner"], sides["x0"]["Sf"], sides["x0"]["Cf"],
                  U=np.tile([u_in, 0.0, 0.0], (n_in, 1)), T=np.full(n_in, T0), rho=np.full(n_in, rho_in))
    if outlet == "zeroGradient":
        out = Patch("outlet", PATCH_ZEROGRADIENT, sides["x1"]["owner"], sides["x1"]["Sf"], sides["x1"]["Cf"])
    else:
        out = Patch("outlet", PATCH_FARFIELD, sides["x1"]["owner"], sides["x1"]["Sf"], sides["x1"]["Cf"],
                    T=np.full(n_out, T0), rho=np.full(n_out, rho0))
    tag = "zg" if outlet == "zeroGradient" else ""
    geo = _assemble(f"openchannel{tag}_{nx}x{ny}_{nv}", d, C, V, own, nei, Sf, Cf, [walls, inlet, out])
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(max(nx / Lx, float(ny)), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, nx=nx, ny=ny, nv=nv, rho_in=rho_in, u_in=u_in))


def pressure_channel(nx: int = 20, ny: int = 8, Lx: float = 2.0, nv: int = 16, kn: float = 0.2,
                     p_ratio: float = 1.2, R: float = 208.0, T0: float = 273.0, rho0: float = 1.0,
                     xi_sigma: float = 4.0, Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0,
                     co: float = 0.5) -> Case:
    """Pressure-driven 2-D channel [0, Lx] x [0, 1]: diffuse walls at y = 0, 1 (T0); a pressureIn
    patch at x = 0 (p_ratio rho0 R T0, T0) and a pressureOut patch at x = Lx (rho0 R T0)."""
    d, C, V, own, nei, Sf, Cf, sides = box_mesh(nx, ny, 0, Lx, 1.0)
    walls = _wall("walls", _concat_sides([sides["y0"], sides["y1"]]), (0, 0, 0), T0)
    n_in, n_out = sides["x0"]["owner"].size, sides["x1"]["owner"].size
    p0 = rho0 * R * T0
    inlet = Patch("inlet", PATCH_PRESSURE_IN, sides["x0"]["owner"], sides["x0"]["Sf"], sides["x0"]["Cf"],
                  T=np.full(n_in, T0), pressure=p_ratio * p0)
    outlet = Patch("outlet", PATCH_PRESSURE_OUT, sides["x1"]["owner"], sides["x1"]["Sf"], sides["x1"]["Cf"],
                   T=np.full(n_out, T0), pressure=p0)
    geo = _assemble(f"pchannel_{nx}x{ny}_{nv}", d, C, V, own, nei, Sf, Cf, [walls, inlet, outlet])
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(max(nx / Lx, float(ny)), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, nx=nx, ny=ny, nv=nv, p_ratio=p_ratio))


def channel(nx: int = 200, ny: int = 40, Lx: float = 5.0, Ly: float = 1.0, nv: int = 32,
            kn: float = 0.5, n_modes: int = 16, amp: float = 0.02, seed: int = 2,
            R: float = 208.0, T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 4.0,
            Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5,
            a: np.ndarray = None) -> Case:
    """Periodic micro-channel: cyclic in x, lower wall T(x) = T0 (1 + sum a_k sin(2 pi k x / Lx)),
    upper wall at T0, gas initially at rest."""
    d, C, V, own, nei, Sf, Cf, sides = box_mesh(nx, ny, 0, Lx, Ly)
    if a is None:
        a = np.random.default_rng(seed).uniform(-amp, amp, n_modes)
    a = np.asarray(a, dtype=np.float64)
    xb = sides["y0"]["Cf"][:, 0]
    Tlow = T0 * (1.0 + np.sum(a[None, :] * np.sin(2 * np.pi * np.arange(1, a.size + 1)[None, :]
                                                     * xb[:, None] / Lx), axis=1))
    lower = _wall("lowerWall", sides["y0"], (0, 0, 0), 0.0)
    lower.T = Tlow
    upper = _wall("upperWall", sides["y1"], (0, 0, 0), T0)
    left = Patch("left", PATCH_CYCLIC, sides["x0"]["owner"], sides["x0"]["Sf"], sides["x0"]["Cf"],
                 neighbour_patch=3)
    right = Patch("right", PATCH_CYCLIC, sides["x1"]["owner"], sides["x1"]["Sf"], sides["x1"]["Cf"],
                  neighbour_patch=2)
    geo = _assemble(f"channel_{nx}x{ny}_{nv}", d, C, V, own, nei, Sf, Cf, [lower, upper, left, right])
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    dt = courant_dt(max(nx / Lx, ny / Ly), d, xi_max, co)
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, a=a, nx=nx, ny=ny, Lx=Lx, T0=T0))


def hard_core_points(n: int, r: float, fixed: np.ndarray, rng, margin: float) -> np.ndarray:
    """n seeded uniform-random points in [margin, 1-margin]^2 at distance >= r from each other
    and from `fixed` (random sequential addition: uniform candidates, rejected when closer
    than r to an accepted point)."""
    h = r / math.sqrt(2.0)
    ng = int(math.ceil(1.0 / h)) + 1
    grid = -np.ones((ng, ng), dtype=np.int64)
    pts = [tuple(p) for p in fixed]
    for i, (x, y) in enumerate(pts):
        grid[min(int(x / h), ng - 1), min(int(y / h), ng - 1)] = i
    out = []
    r2 = r * r
    tries = 0
    while len(out) < n:
        cand = rng.uniform(margin, 1 - margin, size=(4096, 2))
        for x, y in cand:
            tries += 1
            gx, gy = int(x / h), int(y / h)
            ok = True
            for i in range(max(gx - 2, 0), min(gx + 3, ng)):
                for j in range(max(gy - 2, 0), min(gy + 3, ng)):
                    q = grid[i, j]
                    if q >= 0:
                        px, py = pts[q]
                        if (px - x) ** 2 + (py - y) ** 2 < r2:
                            ok = False
                            break
                if not ok:
                    break
            if ok:
                grid[gx, gy] = len(pts)
                pts.append((x, y))
                out.append((x, y))
                if len(out) == n:
                    break
        if tries > 200 * n:
            raise RuntimeError("hard-core sampling did not reach the requested point count")
    return np.array(out)


def fv_courant_dt(c_V, owner, neighbour, Sf, nfi, speed, co):
    """dt for which sum_f |Sf| speed dt / V <= co in every cell (finite-volume CFL)."""
    tot = np.zeros(c_V.shape[0])
    a = np.linalg.norm(Sf, axis=1)
    np.add.at(tot, owner, a)
    np.add.at(tot, neighbour, a[:nfi])
    return float(co * np.min(c_V / (tot * speed)))


def delaunay_cavity(n_interior: int = 99400, n_side: int = 400, nv: int = 64, kn: float = 1.0,
                    seed: int = 0, R: float = 208.0, T0: float = 273.0, rho0: float = 1.0,
                    xi_sigma: float = 5.0, lid_mach: float = 0.16, Pr: float = 2.0 / 3.0,
                    omega: float = 0.5, K: int = 0, co: float = 0.5, dz: float = 0.01,
                    min_sep: float = 0.5, smooth: int = 8, wall_layer: bool = False) -> Case:
    """Cavity on a Delaunay triangulation of seeded uniform-random interior points plus
    n_side uniform points per side of the unit square (one prism layer thick).

    min_sep > 0 keeps interior points at least min_sep / sqrt(n_interior) apart (a
    hard-core random process of uniform density): plain uniform points produce sliver
    triangles on which the scheme is unstable at the reference's Courant number, whose
    deltaCoeffs ignore the cell volume.  `smooth` Laplacian sweeps (each interior point
    moved to the mean of its Delaunay neighbours) then remove the nearly flat triangles
    a hard-core process still allows.  dt is the stricter of that Courant number and the
    finite-volume bound sum_f |Sf| sqrt(2) xi_max dt / V <= 2 co (at most about half of
    a cell's faces are outflow faces for any one velocity).  Cells are renumbered along a Hilbert curve of
    their centroids for locality; faces follow OpenFOAM's upper-triangular order.

    wall_layer: one row of isosceles triangles along the walls, the apex of each above the
    midpoint of its wall edge (height sqrt(3)/2 of the edge), so that every wall cell's centroid
    lies on its wall face's normal through the face centre; the random points fill the rest."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    t = np.arange(n_side) / n_side
    bnd = np.concatenate([np.stack([t, np.zeros_like(t)], 1),
                          np.stack([np.ones_like(t), t], 1),
                          np.stack([1 - t, np.ones_like(t)], 1),
                          np.stack([np.zeros_like(t), 1 - t], 1)])
    fixed = bnd
    lo = 0.0
    if wall_layer:
        # apexes above the wall-edge midpoints, except on the two edges at each corner
        hgt = 0.5 * math.sqrt(3.0) / n_side
        m = (np.arange(1, n_side - 1) + 0.5) / n_side
        lay = np.concatenate([np.stack([m, np.full_like(m, hgt)], 1),
                              np.stack([np.full_like(m, 1 - hgt), m], 1),
                              np.stack([1 - m, np.full_like(m, 1 - hgt)], 1),
                              np.stack([np.full_like(m, hgt), 1 - m], 1)])
        fixed = np.concatenate([bnd, lay])
        lo = hgt
    if min_sep > 0:
        r = min_sep / math.sqrt(n_interior)
        pin = hard_core_points(n_interior, r, fixed, rng, margin=lo + 0.5 * r)
    else:
        margin = lo + 0.25 / n_side
        pin = rng.uniform(margin, 1 - margin, size=(n_interior, 2))
    pts = np.concatenate([fixed, pin])
    nb_ = fixed.shape[0]
    for _ in range(smooth):
        sim = Delaunay(pts).simplices
        e = np.concatenate([sim[:, [0, 1]], sim[:, [1, 2]], sim[:, [2, 0]]])
        e = np.unique(np.sort(e, axis=1), axis=0)
        acc = np.zeros_like(pts)
        cnt = np.zeros(pts.shape[0])
        np.add.at(acc, e[:, 0], pts[e[:, 1]])
        np.add.at(acc, e[:, 1], pts[e[:, 0]])
        np.add.at(cnt, e[:, 0], 1)
        np.add.at(cnt, e[:, 1], 1)
        pts[nb_:] = acc[nb_:] / cnt[nb_:, None]
    tri = Delaunay(pts).simplices.astype(np.int64)
    P = pts[tri]                                   # [nt,3,2]
    area2 = ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1])
             - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))
    keep = np.abs(area2) > 1e-14
    tri, P, area2 = tri[keep], P[keep], area2[keep]
    flip = area2 < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    P = pts[tri]
    area = 0.5 * np.abs(area2)
    cen = P.mean(axis=1)
    order = np.argsort(_hilbert_key(cen[:, 0], cen[:, 1]), kind="stable")
    tri, cen, area = tri[order], cen[order], area[order]
    nt = tri.shape[0]
    # edges
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    cell = np.concatenate([np.arange(nt)] * 3)
    key = np.sort(e, axis=1)
    kid = key[:, 0] * pts.shape[0] + key[:, 1]
    srt = np.argsort(kid, kind="stable")
    kid_s = kid[srt]
    same = np.nonzero(kid_s[1:] == kid_s[:-1])[0]
    a_idx = srt[same]
    b_idx = srt[same + 1]
    ca, cb = cell[a_idx], cell[b_idx]
    own_i = np.minimum(ca, cb)
    nei_i = np.maximum(ca, cb)
    ek = key[a_idx]
    is_pair = np.zeros(e.shape[0], bool)
    is_pair[a_idx] = True
    is_pair[b_idx] = True
    bsel = ~is_pair
    C3 = np.stack([cen[:, 0], cen[:, 1], np.full(nt, 0.5 * dz)], axis=1)
    V = area * dz

    def face_geom(edges, owners):
        p0, p1 = pts[edges[:, 0]], pts[edges[:, 1]]
        t = p1 - p0
        s = np.stack([t[:, 1], -t[:, 0], np.zeros(len(t))], axis=1) * dz
        cf = np.stack([0.5 * (p0[:, 0] + p1[:, 0]), 0.5 * (p0[:, 1] + p1[:, 1]),
                       np.full(len(t), 0.5 * dz)], axis=1)
        return s, cf

    Sf_i, Cf_i = face_geom(ek, own_i)
    sgn = np.sign(np.einsum("ij,ij->i", Sf_i, C3[nei_i] - C3[own_i]))
    Sf_i *= sgn[:, None]
    fo = np.lexsort((nei_i, own_i))
    own_i, nei_i, Sf_i, Cf_i = own_i[fo], nei_i[fo], Sf_i[fo], Cf_i[fo]
    be = e[bsel]
    bo = cell[bsel]
    Sf_b, Cf_b = face_geom(be, bo)
    sgn = np.sign(np.einsum("ij,ij->i", Sf_b, Cf_b - C3[bo]))
    Sf_b *= sgn[:, None]
    top = Cf_b[:, 1] > 1 - 1e-12
    sides_t = dict(owner=bo[top], Sf=Sf_b[top], Cf=Cf_b[top])
    sides_o = dict(owner=bo[~top], Sf=Sf_b[~top], Cf=Cf_b[~top])
    # order boundary faces by owner cell within each patch
    for sd in (sides_t, sides_o):
        o = np.argsort(sd["owner"], kind="stable")
        for kk in sd:
            sd[kk] = sd[kk][o]
    u_lid = lid_mach * math.sqrt(GAMMA_MONATOMIC * R * T0)
    walls = [_wall("movingWall", sides_t, (u_lid, 0, 0), T0),
             _wall("fixedWalls", sides_o, (0, 0, 0), T0)]
    geo = _assemble(f"delaunay_{nt}_{nv}", 2, C3, V, own_i, nei_i, Sf_i, Cf_i, walls)
    xi_max = xi_sigma * math.sqrt(R * T0)
    xi, w = dv_newton_cotes(xi_max, nv)
    mu = mu_from_kn(kn, R, T0, rho0)
    nf = geo["neighbour"].shape[0]
    dcoef = 1.0 / np.linalg.norm(geo["C"][geo["neighbour"]] - geo["C"][geo["owner"][:nf]], axis=1)
    dt = min(courant_dt(float(dcoef.max()), 2, xi_max, co),
             fv_courant_dt(geo["V"], geo["owner"], geo["neighbour"], geo["Sf"], nf, math.sqrt(2) * xi_max, 2 * co))
    return _finish(geo, xi, w, xi_max, R, omega, T0, mu, Pr, K, rho0, T0, dt,
                   meta=dict(kn=kn, seed=seed, n_interior=n_interior, n_side=n_side))


def _hilbert_key(x, y, order: int = 16):
    """Hilbert-curve index of points in the unit square."""
    n = 1 << order
    xi = np.clip((x * n).astype(np.int64), 0, n - 1)
    yi = np.clip((y * n).astype(np.int64), 0, n - 1)
    d = np.zeros_like(xi)
    s = n >> 1
    while s > 0:
        rx = ((xi & s) > 0).astype(np.int64)
        ry = ((yi & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate
        m = ry == 0
        flip = m & (rx == 1)
        xi = np.where(flip, s - 1 - xi, xi)
        yi = np.where(flip, s - 1 - yi, yi)
        xi2 = np.where(m, yi, xi)
        yi2 = np.where(m, xi, yi)
        xi, yi = xi2, yi2
        s >>= 1
    return d


def batch(n_cases: int = 256, nmin: int = 20, nmax: int = 80, nv: int = 28, seed: int = 1,
          kn_range=(0.01, 10.0), lid_range=(0.05, 0.3), tw_range=(250.0, 350.0),
          R: float = 208.0, T0: float = 273.0, rho0: float = 1.0, xi_sigma: float = 4.0,
          Pr: float = 2.0 / 3.0, omega: float = 0.5, K: int = 0, co: float = 0.5,
          indices=None) -> Case:
    """Disjoint union of independent 2-D cavities (side s ~ U{nmin..nmax}, Kn log-uniform,
    lid speed and wall temperature uniform), advanced together with per-case dt and muRef.
    `indices` keeps only those cases of the n_cases drawn (the draws are unchanged)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=n_cases)
    kns = np.exp(rng.uniform(np.log(kn_range[0]), np.log(kn_range[1]), size=n_cases))
    lids = rng.uniform(lid_range[0], lid_range[1], size=n_cases)
    tws = rng.uniform(tw_range[0], tw_range[1], size=n_cases)
    keep = np.arange(n_cases) if indices is None else np.asarray(indices, dtype=np.int64)
    parts = []
    for ci in keep:
        c = cavity(int(sizes[ci]), nv, float(kns[ci]), 2, R, T0, rho0, xi_sigma, float(lids[ci]),
                   Pr, omega, K, co, T_wall=float(tws[ci]))
        parts.append(c)
    name = f"batch{n_cases}_{nv}" if indices is None else f"batch{n_cases}_{nv}_sel{keep.size}"
    return union(parts, name=name,
                 meta=dict(sizes=sizes[keep], kn=kns[keep], lid=lids[keep], T_wall=tws[keep], seed=seed,
                           indices=keep))


def union(parts: List[Case], name: str = "union", meta: dict = None) -> Case:
    """Disjoint union of cases sharing dimension, velocity set and gas constants."""
    p0 = parts[0]
    coff, foff = 0, 0
    C, V, own_i, nei_i, Sf_i, Cf_i, w_i = [], [], [], [], [], [], []
    for p in parts:
        nf = p.n_internal_faces
        C.append(p.C); V.append(p.V)
        own_i.append(p.owner[:nf] + coff); nei_i.append(p.neighbour + coff)
        Sf_i.append(p.Sf[:nf]); Cf_i.append(p.Cf[:nf]); w_i.append(p.weights[:nf])
        coff += p.n_cells
    # boundary: concatenate patch by patch over cases (one patch per case patch)
    pl, own_b, Sf_b, Cf_b, w_b, d_b = [], [], [], [], [], []
    nfi = sum(p.n_internal_faces for p in parts)
    start = nfi
    coff = 0
    pidx_off = 0
    for p in parts:
        nf = p.n_internal_faces
        for pp in p.patches:
            s, n = pp["start"], pp["size"]
            own_b.append(p.owner[s:s + n] + coff)
            Sf_b.append(p.Sf[s:s + n]); Cf_b.append(p.Cf[s:s + n])
            w_b.append(p.weights[s:s + n]); d_b.append(p.delta[s - nf:s - nf + n])
            q = dict(pp)
            q["start"] = start
            if q["neighbour_patch"] >= 0:
                q["neighbour_patch"] += pidx_off
            pl.append(q)
            start += n
        pidx_off += len(p.patches)
        coff += p.n_cells
    cell_case = np.concatenate([np.full(p.n_cells, i, np.int32) for i, p in enumerate(parts)])
    return Case(name=name, dim=p0.dim, C=np.concatenate(C), V=np.concatenate(V),
                owner=np.concatenate(own_i + own_b).astype(np.int32),
                neighbour=np.concatenate(nei_i).astype(np.int32),
                Sf=np.concatenate(Sf_i + Sf_b), Cf=np.concatenate(Cf_i + Cf_b),
                weights=np.concatenate(w_i + w_b), delta=np.concatenate(d_b),
                patches=pl, xi=p0.xi, xi_weights=p0.xi_weights, xi_max=p0.xi_max, R=p0.R,
                omega=p0.omega, T_ref=p0.T_ref, mu_ref=p0.mu_ref, Pr=p0.Pr, K=p0.K,
                rho0=np.concatenate([p.rho0 for p in parts]),
                U0=np.concatenate([p.U0 for p in parts]),
                T0=np.concatenate([p.T0 for p in parts]), dt=min(p.dt for p in parts),
                n_cases=len(parts), cell_case=cell_case,
                mu_ref_case=np.array([p.mu_ref for p in parts]),
                dt_case=np.array([p.dt for p in parts]), meta=meta or {})


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------

def config(name: str, **kw) -> Case:
    """Named workloads: 'cavity2d' (60^2 x 41^2), 'delaunay' (~200k triangles x 64^2),
    'cavity3d' (48^3 x 24^3), 'batch' (256 cavities x 28^2), 'channel' (200x40 x 32^2), and
    'cavity3d_share' (48^3 x 12^3: the mesh of 'cavity3d' with 1728 = 13824 / 8 velocities, the
    velocity count one GPU holds when 'cavity3d' is split over 8 GPUs in velocity space)."""
    if name == "cavity2d":
        return cavity(**{**dict(n=60, nv=41, kn=0.075, dim=2), **kw})
    if name == "delaunay":
        return delaunay_cavity(**{**dict(wall_layer=True), **kw})
    if name == "cavity3d":
        return cavity(**{**dict(n=48, nv=24, kn=0.1, dim=3), **kw})
    if name == "cavity3d_share":
        return cavity(**{**dict(n=48, nv=12, kn=0.1, dim=3), **kw})
    if name == "batch":
        return batch(**kw)
    if name == "channel":
        return channel(**kw)
    raise KeyError(name)
