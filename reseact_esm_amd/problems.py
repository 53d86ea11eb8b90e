"""Synthetic workloads of the operator-split step: grid, wind, mechanism,
initial state, emissions and photolysis factors.

Every array is generated here, on the host, from a seed, and handed to the C
ABI; the GPU library and the CPU oracle therefore see bit-identical inputs.

Definitions (shared by every workload):
  * sphere radius a = 6.371e6 m, model top H = 10 km, NLEV equal layers;
    longitude centres (i+1/2) dlambda (periodic), latitude centres
    -90 deg + (j+1/2) dphi, level centres (k+1/2) H/NLEV, k = 0 at the surface;
  * wind: u = 20 cos(phi) m/s, v = 0, w = w0 sin(pi z/H) cos(lambda) at the
    vertical faces (zero at bottom and top), w0 chosen so that the largest
    vertical Courant number at the reference step equals the target;
  * mechanisms: 1-2 reactants and 1-2 products per reaction, with
    sum(nu_products) <= sum(nu_reactants); rate constants log-uniform;
    photolysis reactions are first order and scale with max(0, cos SZA);
  * cos SZA is the reference's solar formula (run_reseact.jl:164-177) at
    t = 5400 s + model time, evaluated once per split interval (forcing is
    held constant inside a window, rx_traced_integrator.jl:9-10).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

A_EARTH = 6.371e6
H_TOP = 10000.0
U_SOLID = 20.0
K_GAMMA = 0.01721420632103996     # run_reseact.jl:163
D2R = 0.017453292519943295


@dataclasses.dataclass
class Grid:
    nlon: int
    nlat: int
    nlev: int

    @property
    def nc(self) -> int:
        return self.nlon * self.nlat * self.nlev

    @property
    def ncol(self) -> int:
        return self.nlon * self.nlat

    @property
    def dlam(self) -> float:
        return 2.0 * math.pi / self.nlon

    @property
    def dphi(self) -> float:
        return math.pi / self.nlat

    @property
    def dz(self) -> float:
        return H_TOP / self.nlev

    @property
    def lon(self) -> np.ndarray:          # radians, cell centres
        return (np.arange(self.nlon) + 0.5) * self.dlam

    @property
    def lat(self) -> np.ndarray:          # radians, cell centres
        return -0.5 * math.pi + (np.arange(self.nlat) + 0.5) * self.dphi

    def cell_volume_weights(self) -> np.ndarray:
        """Relative cell volume (proportional to cos(phi)), shape (nlev, nlat, nlon)."""
        w = np.cos(self.lat)[None, :, None] * np.ones((self.nlev, 1, self.nlon))
        return w


@dataclasses.dataclass
class Wind:
    cx: np.ndarray       # (nlat,)   1/s
    wz: np.ndarray       # (nlev+1,) 1/s (divided by dz)
    wl: np.ndarray       # (nlon,)
    w0: float            # m/s

    def max_courant(self, grid: Grid, dt: float) -> tuple[float, float]:
        zon = float(np.max(np.abs(self.cx)) * dt)
        ver = float(np.max(np.abs(self.wz)) * np.max(np.abs(self.wl)) * dt)
        return zon, ver


def solid_body_wind(grid: Grid, dt_ref: float, courant: float = 0.8) -> Wind:
    """u = 20 cos(phi), v = 0, w = w0 sin(pi z/H) cos(lambda)."""
    lat = grid.lat
    cx = U_SOLID * np.cos(lat) / (A_EARTH * np.cos(lat) * grid.dlam)
    zf = np.arange(grid.nlev + 1) * grid.dz
    prof = np.sin(np.pi * zf / H_TOP)
    prof[0] = 0.0
    prof[-1] = 0.0
    wl = np.cos(grid.lon)
    pmax = float(np.max(np.abs(prof))) if grid.nlev > 1 else 0.0
    wmax = pmax * float(np.max(np.abs(wl)))
    w0 = courant * grid.dz / (dt_ref * wmax) if wmax > 0 else 0.0
    wz = w0 * prof / grid.dz
    return Wind(cx=np.ascontiguousarray(cx), wz=np.ascontiguousarray(wz),
                wl=np.ascontiguousarray(wl), w0=w0)


def courant_dt(grid: Grid, courant: float = 0.8) -> float:
    """Largest dt with zonal Courant number `courant` (independent of phi)."""
    return courant * A_EARTH * grid.dlam / U_SOLID


@dataclasses.dataclass
class Mechanism:
    ns: int
    reactants: np.ndarray   # (nr, 2) int32, -1 = none
    products: np.ndarray    # (nr, 2) int32, -1 = none
    k: np.ndarray           # (nr,) float64
    photo: np.ndarray       # (nr,) int32

    @property
    def nr(self) -> int:
        return int(self.k.shape[0])

    def net(self) -> list[list[tuple[int, int]]]:
        out = []
        for r in range(self.nr):
            nu: dict[int, int] = {}
            for a in self.reactants[r]:
                if a >= 0:
                    nu[int(a)] = nu.get(int(a), 0) - 1
            for p in self.products[r]:
                if p >= 0:
                    nu[int(p)] = nu.get(int(p), 0) + 1
            out.append(sorted((s, v) for s, v in nu.items() if v != 0))
        return out

    def flops(self) -> dict:
        """Per-attempt flop terms of one ROS23 attempt on one cell."""
        net = self.net()
        f_rhs = 0
        f_jac = 0
        for r in range(self.nr):
            nre = int(np.sum(self.reactants[r] >= 0))
            order_cost = 1 + (1 if nre == 2 else 0) + (1 if self.photo[r] else 0)
            f_rhs += order_cost + 2 * len(net[r])
            f_jac += nre * 2 * len(net[r]) + (2 if nre == 2 else 0)
        return {"rhs": f_rhs, "jac": f_jac}


def random_mechanism(ns: int, nr: int, seed: int, klo: float, khi: float,
                     nphoto: int = 0) -> Mechanism:
    """Synthetic mass-action mechanism; reactions 0..nphoto-1 are photolysis."""
    rng = np.random.default_rng([seed, 1])
    re = -np.ones((nr, 2), dtype=np.int32)
    pr = -np.ones((nr, 2), dtype=np.int32)
    photo = np.zeros(nr, dtype=np.int32)
    for r in range(nr):
        nre = 1 if r < nphoto else int(rng.integers(1, 3))
        a = int(rng.integers(0, ns))
        re[r, 0] = a
        if nre == 2:
            re[r, 1] = int(rng.integers(0, ns))
        npr = 1 if nre == 1 else int(rng.integers(1, 3))
        for q in range(npr):
            p = int(rng.integers(0, ns))
            while nre == 1 and p == a and ns > 1:
                p = int(rng.integers(0, ns))
            pr[r, q] = p
        photo[r] = 1 if r < nphoto else 0
    k = 10.0 ** rng.uniform(math.log10(klo), math.log10(khi), nr)
    return Mechanism(ns=ns, reactants=re, products=pr, k=k, photo=photo)


def cos_sza(t: float, lat_deg: np.ndarray, lon_deg: np.ndarray) -> np.ndarray:
    """Solar zenith cosine, the formula of run_reseact.jl:168-176."""
    g = K_GAMMA * (-12 / 24 + t / 86400)
    dec = (0.006918 - 0.399912 * math.cos(g) + 0.070257 * math.sin(g)
           - 0.006758 * math.cos(2 * g) + 0.000907 * math.sin(2 * g)
           - 0.002697 * math.cos(3 * g) + 0.00148 * math.sin(3 * g))
    eqt = 229.18 * (0.000075 + 0.001868 * math.cos(g) - 0.032077 * math.sin(g)
                    - 0.014615 * math.cos(2 * g) - 0.040849 * math.sin(2 * g))
    w = D2R * ((t / 60 + eqt + 4 * lon_deg) / 4 - 180)
    v = np.sin(D2R * lat_deg) * math.sin(dec) + np.cos(D2R * lat_deg) * math.cos(dec) * np.cos(w)
    return np.clip(v, -1.0, 1.0)


def photolysis_factor(grid: Grid, t_model: float) -> np.ndarray:
    """max(0, cos SZA) per cell (all levels of a column share it), length NC."""
    lat_deg = np.degrees(grid.lat)[:, None]
    lon_deg = np.degrees(grid.lon)[None, :]
    c = np.maximum(cos_sza(5400.0 + t_model, lat_deg, lon_deg), 0.0)   # (nlat, nlon)
    return np.ascontiguousarray(np.broadcast_to(c, (grid.nlev, grid.nlat, grid.nlon)).reshape(-1))


def lognormal_state(grid: Grid, ns: int, seed: int, sigma: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng([seed, 2])
    return rng.lognormal(0.0, sigma, size=ns * grid.nc)


def blob_state(grid: Grid, ns: int, seed: int, nblob: int = 3, amp: float = 5.0) -> np.ndarray:
    """Lognormal background plus Gaussian blobs, species-major."""
    rng = np.random.default_rng([seed, 2])
    u = rng.lognormal(0.0, 1.0, size=(ns, grid.nlev, grid.nlat, grid.nlon))
    lon = grid.lon[None, None, :]
    lat = grid.lat[None, :, None]
    z = ((np.arange(grid.nlev) + 0.5) * grid.dz)[:, None, None]
    for s in range(ns):
        for _ in range(nblob):
            l0 = rng.uniform(0, 2 * math.pi)
            p0 = rng.uniform(-1.2, 1.2)
            z0 = rng.uniform(0.2, 0.8) * H_TOP
            dl = np.angle(np.exp(1j * (lon - l0)))
            r2 = (dl * np.cos(lat)) ** 2 + (lat - p0) ** 2
            blob = np.exp(-r2 / (2 * 0.15 ** 2)) * np.exp(-((z - z0) ** 2) / (2 * (0.15 * H_TOP) ** 2))
            u[s] += amp * blob
    return np.ascontiguousarray(u.reshape(-1))


def lognormal_emissions(grid: Grid, ns: int, seed: int, scale: float = 1e-3,
                        perturb_sigma: float = 0.0, perturb_seed: int | None = None) -> np.ndarray:
    rng = np.random.default_rng([seed, 3])
    e = rng.lognormal(0.0, 1.0, size=ns * grid.ncol) * scale
    if perturb_sigma > 0:
        prng = np.random.default_rng([seed if perturb_seed is None else perturb_seed, 4])
        e = e * prng.lognormal(0.0, perturb_sigma, size=e.shape)
    return e


@dataclasses.dataclass
class Problem:
    """One workload: everything the C ABI needs plus its run parameters."""
    name: str
    grid: Grid
    ns: int
    nt: int
    wind: Wind
    mech: Mechanism | None
    u0: np.ndarray
    jfac: np.ndarray | None = None
    emis: np.ndarray | None = None
    dt_macro: float = 600.0
    nmacro: int = 1
    t0: float = 0.0
    rtol_T: float = 1e-3
    atol_T: float = 1e-6
    rtol_C: float = 1e-3
    atol_C: float = 1e-6
    dt0_T: float = 15.0      # run_reseact_reactant.jl:134
    dt0_C: float = 0.5       # run_reseact_reactant.jl:135
    extra: dict = dataclasses.field(default_factory=dict)
    ppm: "PPMSetup | None" = None   # set: monotone PPM transport (species 0 = air mass)

    @property
    def nc(self) -> int:
        return self.grid.nc

    def tstops(self) -> np.ndarray:
        return self.t0 + self.dt_macro * np.arange(self.nmacro + 1, dtype=np.float64)


def objective_weights(grid: Grid, ns: int, species: int, scope: str) -> tuple[np.ndarray, int]:
    """Weights of the linear objective J = <w, u>: the mean of one species
    over a scope, the reference's objective_weights(var_map, "<prefix>:<scope>")
    (tools/adjoint_gradient.jl:408-426).  scope is "all", "surf" (the lowest
    level, the reference's k == 1) or "levK" (level K counted from 1 at the
    surface, as the reference's state names).  Returns (w, number of states
    selected); w = 1/count on those, 0 elsewhere (species-major, NS*NC)."""
    if not 0 <= species < ns:
        raise ValueError(f"species {species} out of range")
    if scope == "all":
        levels = range(grid.nlev)
    elif scope == "surf":
        levels = [0]
    elif scope.startswith("lev"):
        kk = int(scope[3:])
        if not 1 <= kk <= grid.nlev:
            raise ValueError(f'objective "{scope}" selected no states')
        levels = [kk - 1]
    else:
        raise ValueError(f'unknown objective scope "{scope}"')
    w = np.zeros(ns * grid.nc)
    plane = grid.nlon * grid.nlat
    idx = np.concatenate([species * grid.nc + k * plane + np.arange(plane) for k in levels])
    w[idx] = 1.0 / len(idx)
    return w, len(idx)


def latitude_rows(p: Problem, rows) -> Problem:
    """The workload restricted to a set of latitude rows of its own grid, for
    a bounded CPU sample of the same cells: the same longitudes, levels,
    latitudes' zonal rates, state, emissions and photolysis factors on those
    rows.  With the synthetic wind (v = 0) latitude rows do not interact in
    the transport; the chemistry is per cell.  Only the global error norm sees
    fewer cells.  Upwind-transport workloads only."""
    if p.ppm is not None:
        raise ValueError("latitude_rows: upwind-transport workloads only")
    rows = np.asarray(rows, dtype=np.int64)
    g = p.grid
    sub = Grid(g.nlon, rows.size, g.nlev)
    u0 = p.u0.reshape(p.ns, g.nlev, g.nlat, g.nlon)[:, :, rows, :].reshape(-1).copy()
    emis = None if p.emis is None else \
        p.emis.reshape(p.ns, g.nlat, g.nlon)[:, rows, :].reshape(-1).copy()
    jfac = None if p.jfac is None else \
        p.jfac.reshape(g.nlev, g.nlat, g.nlon)[:, rows, :].reshape(-1).copy()
    wind = Wind(cx=np.ascontiguousarray(p.wind.cx[rows]), wz=p.wind.wz, wl=p.wind.wl,
                w0=p.wind.w0)
    return dataclasses.replace(p, name=f"{p.name}-rows{rows.size}", grid=sub, u0=u0, emis=emis,
                               jfac=jfac, wind=wind, extra=dict(p.extra, rows=rows.tolist()))


def cell_subset(p: Problem, cells) -> Problem:
    """Chemistry-only workload on a set of cells of p (cells are independent
    in per-cell chemistry): their states and photolysis factors as a grid of
    len(cells) x 1 x 1.  Workloads without emissions only (emissions sit on
    level 0 of the grid)."""
    if p.emis is not None:
        raise ValueError("cell_subset: workloads without emissions only")
    cells = np.asarray(cells, dtype=np.int64)
    sub = Grid(cells.size, 1, 1)
    u0 = p.u0.reshape(p.ns, p.nc)[:, cells].reshape(-1).copy()
    jfac = None if p.jfac is None else p.jfac[cells].copy()
    return dataclasses.replace(p, name=f"{p.name}-cells{cells.size}", grid=sub, nt=0, u0=u0,
                               jfac=jfac, wind=solid_body_wind(sub, 1.0, 0.0))


def stratified_cells(g: Grid, di: int = 16, dj: int = 6, dk: int = 6) -> np.ndarray:
    """Every di-th longitude, dj-th latitude and dk-th level (both hemispheres,
    day and night columns), as cell indices c = i + NLON (j + NLAT k)."""
    i = np.arange(0, g.nlon, di)
    j = np.arange(0, g.nlat, dj)
    k = np.arange(0, g.nlev, dk)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return (I + g.nlon * (J + g.nlat * K)).reshape(-1)


def config1(nlon: int = 72, nlat: int = 46, nlev: int = 10, seed: int = 1001,
            nmacro: int = 24) -> Problem:
    """Plain coupled forward: 72x46x10, 4 tracers, 10 species / 20 reactions,
    24 split steps of 600 s, rtol 1e-3, atol 1e-6."""
    g = Grid(nlon, nlat, nlev)
    dt = 600.0
    mech = random_mechanism(10, 20, seed, 1e-6, 1e2)
    return Problem(name="coupled-72x46x10-10sp", grid=g, ns=10, nt=4,
                   wind=solid_body_wind(g, dt, 0.8), mech=mech,
                   u0=lognormal_state(g, 10, seed), dt_macro=dt, nmacro=nmacro)


def config2(nlon: int = 360, nlat: int = 180, nlev: int = 72, nt: int = 20,
            seed: int = 1002, nsteps: int = 60) -> Problem:
    """Transport only: 360x180x72, 20 tracers, fixed dt at Courant 0.8, 60 steps."""
    g = Grid(nlon, nlat, nlev)
    dt = courant_dt(g, 0.8)
    return Problem(name="transport-360x180x72-20tr", grid=g, ns=nt, nt=nt,
                   wind=solid_body_wind(g, dt, 0.8), mech=None,
                   u0=blob_state(g, nt, seed), dt_macro=dt, nmacro=nsteps,
                   extra={"dt_fixed": dt})


def config3(nlon: int = 360, nlat: int = 180, nlev: int = 72, seed: int = 1003) -> Problem:
    """Batched chemistry: 4.67M cells, 50 species / 120 reactions (8 photolysis),
    k log-uniform 1e-8..1e4, adaptive per cell over one 1800 s interval."""
    g = Grid(nlon, nlat, nlev)
    mech = random_mechanism(50, 120, seed, 1e-8, 1e4, nphoto=8)
    return Problem(name="chem-360x180x72-50sp", grid=g, ns=50, nt=0,
                   wind=solid_body_wind(g, 1800.0, 0.8), mech=mech,
                   u0=lognormal_state(g, 50, seed), jfac=photolysis_factor(g, 0.0),
                   dt_macro=1800.0, nmacro=1)


def config4(nlon: int = 144, nlat: int = 91, nlev: int = 47, seed: int = 1004,
            nmacro: int = 12) -> Problem:
    """Coupled forward + adjoint: 144x91x47, 20 species / 40 reactions,
    12 split steps of 900 s, surface emissions, loss sum_surface u0^2."""
    g = Grid(nlon, nlat, nlev)
    dt = 900.0
    mech = random_mechanism(20, 40, seed, 1e-6, 1e2)
    return Problem(name="coupled-adjoint-144x91x47-20sp", grid=g, ns=20, nt=20,
                   wind=solid_body_wind(g, dt, 0.8), mech=mech,
                   u0=lognormal_state(g, 20, seed),
                   emis=lognormal_emissions(g, 20, seed), dt_macro=dt, nmacro=nmacro)


def config5_domains(ndom: int = 32, seed: int = 1005) -> list[tuple[int, int, int]]:
    rng = np.random.default_rng([seed, 5])
    return [(int(rng.integers(64, 257)), int(rng.integers(64, 193)), int(rng.integers(20, 48)))
            for _ in range(ndom)]


def config5(ndom: int = 32, seed: int = 1005, nmacro: int = 6,
            sizes: list[tuple[int, int, int]] | None = None) -> list[Problem]:
    """Ensemble: regional domains 64x64x20 .. 256x192x47, 20-species mechanism,
    emissions perturbed lognormally (sigma 0.3) per domain, 6 split steps."""
    sizes = config5_domains(ndom, seed) if sizes is None else sizes
    mech = random_mechanism(20, 40, 1004, 1e-6, 1e2)
    out = []
    for d, (nx, ny, nz) in enumerate(sizes):
        g = Grid(nx, ny, nz)
        out.append(Problem(name=f"ensemble-{d}-{nx}x{ny}x{nz}", grid=g, ns=20, nt=20,
                           wind=solid_body_wind(g, 900.0, 0.8), mech=mech,
                           u0=lognormal_state(g, 20, seed + d),
                           emis=lognormal_emissions(g, 20, 1004, perturb_sigma=0.3,
                                                    perturb_seed=seed + d),
                           dt_macro=900.0, nmacro=nmacro))
    return out


def toy_problem() -> Problem:
    """3 species x 5 cells: A+B -> C, C -> A+B, 2C -> B (the shape of the
    reference's model-free adjoint toy, tools/rx_adjoint_toy.jl:63-74, with
    the time factor s(t) fixed)."""
    g = Grid(5, 1, 1)
    re = np.array([[0, 1], [2, -1], [2, 2]], dtype=np.int32)
    pr = np.array([[2, -1], [0, 1], [1, -1]], dtype=np.int32)
    mech = Mechanism(ns=3, reactants=re, products=pr, k=np.array([0.7, 0.3, 0.15]),
                     photo=np.zeros(3, dtype=np.int32))
    rng = np.random.default_rng(11)
    u0 = 0.5 + rng.random(15)
    return Problem(name="toy-3sp-5cells", grid=g, ns=3, nt=0,
                   wind=solid_body_wind(g, 1.0, 0.0), mech=mech, u0=u0)


# ---------------------------------------------------------------------------
# Monotone PPM transport in CWC mixing-ratio form (the reference's transport,
# reseact.esm:1536-3100): geometry, air-mass fluxes, inflow halos.
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class PPMSetup:
    lon_periodic: bool
    dlon: float                 # rad
    coslat_e: np.ndarray        # (nlat+1,)
    dS_lat: np.ndarray          # (nlat,)
    dphi_lat: np.ndarray        # (nlat,)
    Mx: np.ndarray              # (nlev, nlat, nlon+1)  Pa/s
    My: np.ndarray              # (nlev, nlat+1, nlon)
    Mz: np.ndarray              # (nlev+1, nlat, nlon)
    dp: np.ndarray              # (nlev, nlat, nlon)    Pa
    qbc_w: np.ndarray | None = None   # (nq, nlev, nlat)
    qbc_e: np.ndarray | None = None
    qbc_s: np.ndarray | None = None   # (nq, nlev, nlon)
    qbc_n: np.ndarray | None = None
    dPSdt: np.ndarray | None = None   # (nlat, nlon): column pressure fixer of the air mass
    pbl_f: np.ndarray | None = None   # (nlev, nlat, nlon): layer fraction inside the PBL
    tau_pbl: float = 900.0            # s, PBL mixing relaxation time


G_ACC = 9.80665        # m/s^2 (reseact.esm:1164)
RD_AIR = 287.05        # J/(kg K) (reseact.esm:1158)


def pbl_fraction(dp: np.ndarray, pblh: np.ndarray, temp: float | np.ndarray = 290.0,
                 p_lid: float = 0.0) -> np.ndarray:
    """Fraction of each layer inside the PBL, the reference's F_OF_PBL
    (reseact.esm:7455 p_pbl_top = PS exp(-PBLH g / (Rd T)); :7536 pbl_f =
    clamp((min(PS, p_bot) - max(p_pbl_top, p_top)) / dp, 0, 1)), for layers
    dp (nlev, nlat, nlon) stacked from the surface (level 0) on a lid at
    p_lid: PS = p_lid + sum dp, p_bot/p_top the layer's edge pressures.
    pblh (nlat, nlon) in m, temp the surface-layer temperature."""
    dp = np.asarray(dp, np.float64)
    ps = p_lid + dp.sum(axis=0)
    p_bot = ps[None] - np.concatenate([np.zeros_like(ps)[None], np.cumsum(dp, axis=0)[:-1]])
    p_top = p_bot - dp
    p_pbl = ps * np.exp(-(np.asarray(pblh, np.float64) * G_ACC) / (RD_AIR * np.asarray(temp, np.float64)))
    f = (np.minimum(ps[None], p_bot) - np.maximum(p_pbl[None], p_top)) / dp
    return np.clip(f, 0.0, 1.0)


def with_pbl_mixing(p: Problem, pblh_mean: float = 1000.0, pblh_amp: float = 600.0,
                    tau: float = 900.0) -> Problem:
    """The PPM workload with PBL mixing: a boundary layer of pblh_mean +-
    pblh_amp metres varying with longitude (a day side and a night side)."""
    g = p.grid
    lam = 2.0 * math.pi * (np.arange(g.nlon) + 0.5) / g.nlon
    pblh = np.broadcast_to(pblh_mean + pblh_amp * np.cos(lam)[None, :], (g.nlat, g.nlon))
    p.ppm.pbl_f = pbl_fraction(p.ppm.dp, pblh, p_lid=5000.0)
    p.ppm.tau_pbl = tau
    return p


# GEOS-FP hybrid coefficients at the lowest 8 level edges (surface first),
# prototypes/transport_3d/gen_t3d.py:8-9.
GEOSFP_AP8 = [0.0, 4.804826, 659.3752, 1313.48, 1961.311, 2609.201, 3257.081, 3898.201]
GEOSFP_BP8 = [1.0, 0.984952, 0.963406, 0.941865, 0.920387, 0.898908, 0.877429, 0.856018]


def lat_edges_deg(lat_points_deg: np.ndarray) -> np.ndarray:
    """Cell edges of POINT-based latitude rows: midpoints between points,
    half a spacing beyond the end points, clamped to +-90 (the half-polar
    caps of a global grid), reseact.esm:146-153."""
    d = float(lat_points_deg[1] - lat_points_deg[0]) if lat_points_deg.size > 1 else 1.0
    e = np.concatenate([[lat_points_deg[0] - d / 2], 0.5 * (lat_points_deg[1:] + lat_points_deg[:-1]),
                        [lat_points_deg[-1] + d / 2]])
    return np.clip(e, -90.0, 90.0)


def sphere_metrics(phie_deg: np.ndarray):
    """coslat_e, dS_lat = sin(phie[j+1]) - sin(phie[j]), dphi_lat (rad);
    a face at a pole has length exactly 0."""
    phie = phie_deg * D2R
    ce = np.cos(phie)
    ce[np.abs(phie_deg) == 90.0] = 0.0
    dS = np.sin(phie[1:]) - np.sin(phie[:-1])
    dphi = phie[1:] - phie[:-1]
    return ce, dS, dphi


def transport3d_regional(nlon: int = 7, nlat: int = 7, nlev: int = 7, nq: int = 1,
                         qbc_scale: float = 1.0) -> Problem:
    """The reference's CWC gate case (prototypes/transport_3d/transport_3d.esm,
    test freestream_q_stays_one_regional_inflow): a CONUS patch of GEOS-FP 4x5
    cells (west edge -112.5 deg, 5 deg; latitude points from 26 deg, 4 deg),
    open walls with inflow W and S, outflow E and N; analytic air-mass fluxes
    Mx = 0.7 cos(lon_face)(1 + 0.3 sin(lat_c))(1 + 0.05 k),
    My = 0.5 sin(2 lon_c) cos(lat_face)(1 + 0.05 k),
    Mz = 0.05*4 s(1-s)(1-2s), s = k/NLEV (zero at surface and lid);
    dp from the GEOS-FP hybrid table at PS = 101325 Pa; m(0) = 2 + 0.5 sin(20 +
    5 j deg) + 0.1 cos(5 (i + 1/2) deg) and every mixing ratio 1, with the
    inflow halos at `qbc_scale` (1: the gate; 1.000001: its falsifier)."""
    g = Grid(nlon, nlat, nlev)
    lon0, dlon_deg, lat0, dlat_deg = -112.5, 5.0, 26.0, 4.0
    phie = lat_edges_deg(lat0 + dlat_deg * np.arange(nlat))
    ce, dS, dphi = sphere_metrics(phie)
    k = np.arange(nlev)[:, None, None]
    lon_face = (lon0 + dlon_deg * np.arange(nlon + 1))[None, None, :]
    lat_c = (0.5 * (phie[:-1] + phie[1:]))[None, :, None]
    Mx = 0.7 * np.cos(D2R * lon_face) * (1.0 + 0.3 * np.sin(D2R * lat_c)) * (1.0 + 0.05 * k)
    lon_c = (lon0 + dlon_deg * (np.arange(nlon) + 0.5))[None, None, :]
    My = 0.5 * np.sin(2.0 * D2R * lon_c) * np.cos(D2R * phie[None, :, None]) * (1.0 + 0.05 * k)
    sf = np.arange(nlev + 1) / float(nlev)
    mz = 0.05 * (4.0 * sf * (1.0 - sf) * (1.0 - 2.0 * sf))
    mz[0] = mz[-1] = 0.0
    Mz = np.broadcast_to(mz[:, None, None], (nlev + 1, nlat, nlon))
    ap = np.interp(np.linspace(0, 7, nlev + 1), np.arange(8), GEOSFP_AP8)
    bp = np.interp(np.linspace(0, 7, nlev + 1), np.arange(8), GEOSFP_BP8)
    dpk = (ap[:-1] - ap[1:]) + (bp[:-1] - bp[1:]) * 101325.0
    dp = np.broadcast_to(dpk[:, None, None], (nlev, nlat, nlon))
    jj = np.arange(nlat)[None, :, None]
    ii = np.arange(nlon)[None, None, :]
    m0 = np.broadcast_to(2.0 + 0.5 * np.sin(D2R * (20.0 + jj * 5.0)) + 0.1 * np.cos(D2R * ((ii + 0.5) * 5.0)),
                         (nlev, nlat, nlon))
    ns = 1 + nq
    u0 = np.concatenate([np.ascontiguousarray(m0).reshape(-1), np.ones(nq * g.nc)])
    halo_we = np.full((nq, nlev, nlat), qbc_scale)
    halo_sn = np.full((nq, nlev, nlon), qbc_scale)
    ppm = PPMSetup(False, dlon_deg * D2R, ce, dS, dphi, np.ascontiguousarray(Mx),
                   np.ascontiguousarray(My), np.ascontiguousarray(Mz), np.ascontiguousarray(dp),
                   halo_we, halo_we.copy(), halo_sn, halo_sn.copy())
    return Problem(name=f"transport3d-regional-{nlon}x{nlat}x{nlev}", grid=g, ns=ns, nt=ns,
                   wind=solid_body_wind(g, 1.0, 0.0), mech=None, u0=u0, ppm=ppm,
                   dt_macro=0.1, nmacro=1)


def continuity_mz(g: Grid, ppm_geom, Mx: np.ndarray, My: np.ndarray, dp: np.ndarray) -> np.ndarray:
    """Vertical air-mass flux diagnosed from continuity with the column
    pressure fixer at dPS/dt = 0 (reseact.esm:262-291): divh_fix(k) = divh(k)
    - divh_col dp(k)/dp_col, Mz[k] = -sum_{k'<k} divh_fix(k'), zero at the
    surface and (set exactly) at the lid."""
    dlon, ce, dS, dphi = ppm_geom
    gx = (dphi / (dlon * dS))[None, :, None]
    divh = (Mx[:, :, 1:] - Mx[:, :, :-1]) * gx \
        + (ce[None, 1:, None] * My[:, 1:, :] - ce[None, :-1, None] * My[:, :-1, :]) / dS[None, :, None]
    col = divh.sum(0)
    fix = divh - col[None] * dp / dp.sum(0)[None]
    Mz = np.zeros((g.nlev + 1, g.nlat, g.nlon))
    Mz[1:] = -np.cumsum(fix, axis=0)
    Mz[-1] = 0.0
    return Mz


def ppm_global(nlon: int = 360, nlat: int = 180, nlev: int = 72, nq: int = 20,
               seed: int = 1006) -> Problem:
    """Global PPM transport: periodic longitude, poles (zero-length faces),
    hybrid levels from a smooth surface-to-top pressure profile, air-mass
    fluxes of u = 20 cos(phi) + 5 sin(2 lambda) cos(phi) m/s and
    v = 5 sin(lambda) cos(phi)^2 m/s, Mz from continuity with the pressure
    fixer at dPS/dt = 0 and the same fixer on the air-mass equation (so m is
    stationary to round-off), mixing ratios lognormal.
    Species 0 is the air mass; NT = NS = 1 + nq."""
    g = Grid(nlon, nlat, nlev)
    phie = np.linspace(-90.0, 90.0, nlat + 1)
    ce, dS, dphi = sphere_metrics(phie)
    dlon = 2.0 * math.pi / nlon
    pe = 101325.0 * np.exp(-np.linspace(0.0, 3.0, nlev + 1))       # level edges, Pa
    dpk = pe[:-1] - pe[1:]
    dp = np.ascontiguousarray(np.broadcast_to(dpk[:, None, None], (nlev, nlat, nlon)))
    lam_f = (dlon * np.arange(nlon + 1))[None, None, :]
    lam_c = (dlon * (np.arange(nlon) + 0.5))[None, None, :]
    phc = (0.5 * (phie[:-1] + phie[1:]) * D2R)[None, :, None]
    phf = (phie * D2R)[None, :, None]
    dpx = dpk[:, None, None]
    U = (20.0 + 5.0 * np.sin(2.0 * lam_f)) * np.cos(phc)
    V = 5.0 * np.sin(lam_c) * np.cos(phf) ** 2
    Mx = np.ascontiguousarray(dpx * U / A_EARTH * np.ones((nlev, nlat, 1)))
    Mx[:, :, -1] = Mx[:, :, 0]
    My = np.ascontiguousarray(dpx * V / A_EARTH * np.ones((nlev, 1, nlon)))
    Mz = continuity_mz(g, (dlon, ce, dS, dphi), Mx, My, dp)
    rng = np.random.default_rng([seed, 2])
    q = rng.lognormal(0.0, 0.5, size=nq * g.nc)
    u0 = np.concatenate([dp.reshape(-1), q])
    ppm = PPMSetup(True, dlon, ce, dS, dphi, Mx, My, np.ascontiguousarray(Mz), dp,
                   dPSdt=np.zeros((nlat, nlon)))
    return Problem(name=f"transport-ppm-{nlon}x{nlat}x{nlev}-{nq}tr", grid=g, ns=1 + nq, nt=1 + nq,
                   wind=solid_body_wind(g, 1.0, 0.0), mech=None, u0=u0, ppm=ppm,
                   dt_macro=600.0, nmacro=1)


def ppm_coupled_regional(nlon: int = 10, nlat: int = 8, nlev: int = 6, nq: int = 3,
                         seed: int = 5, nmacro: int = 2) -> Problem:
    """The regional PPM patch (open walls) with lognormal mixing ratios and
    inflow halos, coupled to a small mass-action mechanism among the mixing
    ratios (q1 -> q2, q2 + q3 -> q1, q3 -> q2; the air mass, species 0, takes
    no part): the workload of the discrete adjoint through the PPM transport."""
    if nq < 3:
        raise ValueError("the mechanism needs three mixing ratios")
    p = transport3d_regional(nlon, nlat, nlev, nq=nq)
    rng = np.random.default_rng([seed, 3])
    nc = p.nc
    p.u0 = p.u0.copy()
    p.u0[nc:] = rng.lognormal(0.0, 0.5, nq * nc)
    pp = p.ppm
    for name in ("qbc_w", "qbc_e", "qbc_s", "qbc_n"):
        setattr(pp, name, rng.lognormal(0.0, 0.5, getattr(pp, name).shape))
    re = np.array([[1, -1], [2, 3], [3, -1]], dtype=np.int32)
    pr = np.array([[2, -1], [1, -1], [2, -1]], dtype=np.int32)
    p.mech = Mechanism(ns=1 + nq, reactants=re, products=pr, k=np.array([2.0, 0.5, 1.0]),
                       photo=np.zeros(3, dtype=np.int32))
    p.name = f"ppm-coupled-regional-{nlon}x{nlat}x{nlev}-{nq}q"
    p.dt_macro = 0.05
    p.nmacro = nmacro
    p.dt0_T = 0.01
    p.dt0_C = 0.01
    return p
