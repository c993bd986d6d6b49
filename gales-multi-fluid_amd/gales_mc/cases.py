"""Synthetic meshes, seeded fields and physics parameters of the benchmark cases.

The reference runs pre-partitioned text meshes of a volcanic conduit
(sim/fluid_mc/P0/setup.txt:1) with magma-table properties; the cases here are
the synthetic channels of the benchmark configurations:

  c0  2D channel [0,2]x[0,1], 64x32 quads -> 4096 triangles, 2 components
  c1  2D channel, 2048x1024 quads (interior nodes jittered by U(-0.2h, 0.2h))
  c2  3D box [0,1]^3, 128^3 hexes -> 6 Kuhn tets each, 3 components
  c3  = c1 mesh and physics (adjoint case)
  c4  3D box [0,2]x[0,2]x[0,1], 192x192x96 hexes, 4 components

Boundary semantics follow the reference case pattern (fluid_ic_bc.hpp:90-217):
node flag 5 = wall (velocity Dirichlet), flag 2 = inlet (every dof but p
Dirichlet), flag 3 = outlet (natural); side flag 5 = wall with q = 0 and
J_1 = 0 Neumann overrides, sides 2/3 carry no override.  Corner nodes take the
wall flag.
"""
import numpy as np

from . import abi

WALL, INLET, OUTLET = 5, 2, 3


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def channel2d(nx, ny, Lx=2.0, Ly=1.0, jitter=0.0, seed=0):
    """Structured quads split by the (i,j)-(i+1,j+1) diagonal into two
    anticlockwise triangles.  Returns a dict of numpy arrays."""
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    ii = ii.ravel()
    jj = jj.ravel()
    hx, hy = Lx / nx, Ly / ny
    x = ii * hx
    y = jj * hy
    flag = np.zeros(ii.size, np.int32)
    flag[ii == nx] = OUTLET
    flag[ii == 0] = INLET
    flag[(jj == 0) | (jj == ny)] = WALL
    if jitter > 0.0:
        rng = _rng(seed + 7)
        interior = flag == 0
        x = x + np.where(interior, rng.uniform(-jitter, jitter, ii.size) * hx, 0.0)
        y = y + np.where(interior, rng.uniform(-jitter, jitter, ii.size) * hy, 0.0)
    coords = np.stack([x, y], 1).astype(np.float64)
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny))
    qi = qi.ravel()
    qj = qj.ravel()
    n00 = qj * (nx + 1) + qi
    n10 = n00 + 1
    n01 = n00 + nx + 1
    n11 = n01 + 1
    conn = np.empty((2 * nx * ny, 3), np.int32)
    conn[0::2] = np.stack([n00, n10, n11], 1)   # tri A
    conn[1::2] = np.stack([n00, n11, n01], 1)   # tri B
    q = np.arange(nx * ny)
    se, sl, sf = [], [], []
    bot = q[qj == 0]
    se.append(2 * bot); sl.append(np.full(bot.size, 2)); sf.append(np.full(bot.size, WALL))
    top = q[qj == ny - 1]
    se.append(2 * top + 1); sl.append(np.full(top.size, 0)); sf.append(np.full(top.size, WALL))
    left = q[qi == 0]
    se.append(2 * left + 1); sl.append(np.full(left.size, 1)); sf.append(np.full(left.size, INLET))
    right = q[qi == nx - 1]
    se.append(2 * right); sl.append(np.full(right.size, 0)); sf.append(np.full(right.size, OUTLET))
    return dict(dim=2, coords=coords, conn=conn, node_flag=flag,
                side_elem=np.concatenate(se).astype(np.int32),
                side_local=np.concatenate(sl).astype(np.int32),
                side_flag=np.concatenate(sf).astype(np.int32),
                shape=(nx, ny), extent=(Lx, Ly))


# tet faces in element.hpp:455-459 numbering
TET_FACES = [(1, 2, 3), (0, 3, 2), (0, 1, 3), (0, 2, 1)]


def kuhn3d(nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0):
    """Hexahedra split into 6 tetrahedra along the main diagonal (Kuhn),
    node order fixed so every tet has a positive Jacobian."""
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    ii = ii.transpose(2, 1, 0).ravel()
    jj = jj.transpose(2, 1, 0).ravel()
    kk = kk.transpose(2, 1, 0).ravel()
    coords = np.stack([ii * (Lx / nx), jj * (Ly / ny), kk * (Lz / nz)], 1).astype(np.float64)
    flag = np.zeros(ii.size, np.int32)
    flag[ii == nx] = OUTLET
    flag[ii == 0] = INLET
    flag[(jj == 0) | (jj == ny) | (kk == 0) | (kk == nz)] = WALL
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).ravel()
    cj = cj.transpose(2, 1, 0).ravel()
    ck = ck.transpose(2, 1, 0).ravel()
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    base = (ck * (ny + 1) + cj) * (nx + 1) + ci
    step = [sx, sy, sz]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    parity = [1, -1, -1, 1, 1, -1]
    tets = []
    for (a, b, c), par in zip(perms, parity):
        v0 = base
        v1 = base + step[a]
        v2 = v1 + step[b]
        v3 = v2 + step[c]
        if par > 0:
            tets.append(np.stack([v0, v1, v2, v3], 1))
        else:
            tets.append(np.stack([v0, v1, v3, v2], 1))
    conn = np.stack(tets, 1).reshape(-1, 4).astype(np.int32)
    # boundary faces: all three face nodes on one domain plane
    se, sl, sf = [], [], []
    I3 = np.stack([ii, jj, kk], 1)
    lim = np.array([nx, ny, nz])
    for s, face in enumerate(TET_FACES):
        fn = conn[:, face]                       # (E, 3)
        fc = I3[fn]                              # (E, 3, 3) integer coords
        for ax in range(3):
            for side, val in ((0, 0), (1, lim[ax])):
                on = np.all(fc[:, :, ax] == val, axis=1)
                e = np.nonzero(on)[0]
                if e.size == 0:
                    continue
                if ax == 0:
                    f = INLET if side == 0 else OUTLET
                else:
                    f = WALL
                se.append(e)
                sl.append(np.full(e.size, s))
                sf.append(np.full(e.size, f))
    return dict(dim=3, coords=coords, conn=conn, node_flag=flag,
                side_elem=np.concatenate(se).astype(np.int32),
                side_local=np.concatenate(sl).astype(np.int32),
                side_flag=np.concatenate(sf).astype(np.int32),
                shape=(nx, ny, nz), extent=(Lx, Ly, Lz))


# ---------------------------------------------------------------------------
# physics
# ---------------------------------------------------------------------------
def custom(rho, mu, cp, kappa, alpha=0.0, beta=0.0, D=1e-10):
    return dict(kind=abi.GMC_COMP_CUSTOM, rho=rho, mu=mu, cp=cp, kappa=kappa, alpha=alpha,
                beta=beta, molar_mass=0.0, gamma=0.0, diffusivity=D)


def ideal_gas(molar_mass, gamma=1.4, mu=1e-5, D=1e-10):
    return dict(kind=abi.GMC_COMP_IDEAL_GAS, rho=0.0, mu=mu, cp=0.0, kappa=0.0, alpha=0.0,
                beta=0.0, molar_mass=molar_mass, gamma=gamma, diffusivity=D)


MELT = custom(2400.0, 1e2, 1200.0, 1.5)              # kappa: fluid_properties_reader.hpp:297
CRYSTAL = custom(2800.0, 1e4, 1000.0, 2.0)
CRYSTAL2 = custom(3000.0, 1e5, 900.0, 2.5)
GAS_R461 = ideal_gas(8.314462 / 461.0, 1.4)           # R = 461 J/kg/K
H2O = ideal_gas(0.01802, 1.4)                         # chemicals.hpp:203-207


# magma_mix blocks of the reference's shipped case (sim/fluid_mc/P0/props.txt:
# oxide mass fractions SiO2 TiO2 Al2O3 Fe2O3 FeO MnO MgO CaO Na2O K2O)
SHOSHONITE_OXIDES = (0.5247, 0.0085, 0.1760, 0.0188, 0.0574, 0.0012, 0.0360, 0.0793, 0.0343, 0.0428)
PHONOLITE_OXIDES = (0.5352, 0.0060, 0.1984, 0.0160, 0.0320, 0.0014, 0.0176, 0.0676, 0.0466, 0.0791)


def magma(oxide_wf, h2o_wf, co2_wf, p_table, wf_h2o_l, wf_co2_l, T_table=None,
          gas_on_mu=abi.GMC_MU_ISHII_ZUBER, crystal_on_mu=abi.GMC_MU_NONE, crystal_vf=0.0,
          crystal_rho=0.0, melt_kappa=1.5, h2o_l_cp=2278.0, h2o_l_kappa=0.64, co2_l_cp=2278.0,
          co2_l_kappa=0.64):
    """A magma submixture component (defaults of fluid_properties_reader.hpp:
    293-300).  The tables are numpy arrays: wf_* of shape [n_p], or [n_T, n_p]
    with T_table."""
    tab = dict(p_table=np.ascontiguousarray(p_table, np.float64),
               wf_h2o_l=np.ascontiguousarray(wf_h2o_l, np.float64),
               wf_co2_l=np.ascontiguousarray(wf_co2_l, np.float64))
    if T_table is not None:
        tab["T_table"] = np.ascontiguousarray(T_table, np.float64)
    return dict(kind=abi.GMC_COMP_MAGMA, rho=0.0, mu=0.0, cp=0.0, kappa=0.0, alpha=0.0, beta=0.0,
                molar_mass=0.0, gamma=0.0, diffusivity=1e-10,
                magma=dict(oxide_wf=tuple(oxide_wf), melt_kappa=melt_kappa, h2o_wf=h2o_wf,
                           co2_wf=co2_wf, h2o_l_cp=h2o_l_cp, h2o_l_kappa=h2o_l_kappa,
                           co2_l_cp=co2_l_cp, co2_l_kappa=co2_l_kappa, crystal_vf=crystal_vf,
                           crystal_rho=crystal_rho, gas_on_mu=gas_on_mu,
                           crystal_on_mu=crystal_on_mu, tables=tab))


def _set_magma(m, d, keep):
    import ctypes as C
    for k, v in d.items():
        if k == "tables":
            continue
        if k == "oxide_wf":
            for i in range(10):
                m.oxide_wf[i] = v[i]
        else:
            setattr(m, k, v)
    tab = d["tables"]
    keep.append(tab)
    dp = C.POINTER(C.c_double)
    m.n_p = tab["p_table"].size
    m.n_T = tab["T_table"].size if "T_table" in tab else 0
    m.p_table = tab["p_table"].ctypes.data_as(dp)
    m.T_table = tab["T_table"].ctypes.data_as(dp) if "T_table" in tab else None
    m.wf_h2o_l = tab["wf_h2o_l"].ctypes.data_as(dp)
    m.wf_co2_l = tab["wf_co2_l"].ctypes.data_as(dp)


def make_params(dim, comps, dt=1e-3, gravity=None, dc=True, precision=0):
    """abi.Params of a case.  Magma tables stay referenced from p._keep while
    p lives (gmc_create copies them)."""
    p = abi.Params()
    p.dim = dim
    p.nc = len(comps)
    keep = []
    for c, d in enumerate(comps):
        for k, v in d.items():
            if k == "magma":
                _set_magma(p.comp[c].magma, v, keep)
            else:
                setattr(p.comp[c], k, v)
    p._keep = keep
    g = gravity if gravity is not None else ((0.0, -9.81, 0.0) if dim == 2 else (0.0, 0.0, -9.81))
    for k in range(3):
        p.gravity[k] = g[k]
    p.p_ref = 1e7
    p.T_ref = 1300.0
    p.dt = dt
    p.steady_state = 0
    p.tau_variant = abi.GMC_TAU_NON_DIAG_COMP_2019
    p.dc_2006 = 1 if dc else 0
    p.dc_sharp = 2.0
    p.dc_scale_fact = 1.0
    ndof = dim + 2 + p.nc - 1
    vel = sum(1 << (1 + k) for k in range(dim))
    p.dirichlet_mask[WALL] = vel
    p.dirichlet_mask[INLET] = ((1 << ndof) - 1) & ~1
    nb = p.neumann[WALL]
    nb.q_mask = (1 << dim) - 1
    for k in range(dim):
        nb.J_mask[k] = 1
    p.precision = precision
    return p


# ---------------------------------------------------------------------------
# seeded fields
# ---------------------------------------------------------------------------
def _sines(rng, X, ext, amp=0.25, nterm=4):
    out = np.zeros(X.shape[0])
    for _ in range(nterm):
        kv = rng.integers(1, 4, size=X.shape[1])
        ph = rng.uniform(0.0, 2.0 * np.pi)
        arg = sum(kv[d] * X[:, d] / ext[d] for d in range(X.shape[1]))
        out += amp * np.sin(2.0 * np.pi * arg + ph)
    return out


def fields(mesh, nc, seed):
    """Returns (I, P) node-interleaved FP64 states [n_nodes, ndof].

    P: p = 1e7 + N(0, 1e4); each velocity component a sum of 4 random-phase
    sines of amplitude 1/4; T = 1300 + N(0, 5); mass fractions Y_1 ~ U(0.2, 0.8)
    for 2 components, else a Dirichlet(alpha(x)) draw with alpha_c =
    2(1 + 0.5 sin(2 pi k_c x + phi_c)).  Walls get zero velocity.
    I = P + delta: dp ~ N(0, 10) Pa, dv ~ N(0, 1e-3) m/s (0 on walls),
    dT ~ N(0, 1e-2) K, dY ~ N(0, 1e-3)."""
    X = mesh["coords"]
    dim = mesh["dim"]
    ext = mesh["extent"]
    n = X.shape[0]
    ndof = dim + 2 + nc - 1
    rng = _rng(seed)
    P = np.empty((n, ndof))
    P[:, 0] = 1e7 + rng.normal(0.0, 1e4, n)
    for k in range(dim):
        P[:, 1 + k] = _sines(rng, X, ext)
    P[:, dim + 1] = 1300.0 + rng.normal(0.0, 5.0, n)
    if nc == 2:
        P[:, dim + 2] = rng.uniform(0.2, 0.8, n)
    else:
        alpha = np.empty((n, nc))
        for c in range(nc):
            kc = rng.integers(1, 4)
            ph = rng.uniform(0.0, 2.0 * np.pi)
            alpha[:, c] = 2.0 * (1.0 + 0.5 * np.sin(2.0 * np.pi * kc * X[:, 0] / ext[0] + ph))
        gm = rng.standard_gamma(alpha)
        Y = gm / gm.sum(1, keepdims=True)
        P[:, dim + 2:] = Y[:, : nc - 1]
    wall = mesh["node_flag"] == WALL
    P[wall, 1:dim + 1] = 0.0
    I = P.copy()
    I[:, 0] += rng.normal(0.0, 10.0, n)
    I[:, 1:dim + 1] += np.where(wall[:, None], 0.0, rng.normal(0.0, 1e-3, (n, dim)))
    I[:, dim + 1] += rng.normal(0.0, 1e-2, n)
    I[:, dim + 2:] += rng.normal(0.0, 1e-3, (n, nc - 1))
    return np.ascontiguousarray(I), np.ascontiguousarray(P)


SEED0 = 20261203


def config(cid, scale=1):
    """(mesh, params, I, P) of benchmark configuration cid (0..4).  scale > 1
    divides the mesh resolution (for tests that run the oracle)."""
    if cid in (0, 1, 3):
        if cid == 0:
            nx, ny, jit = 64, 32, 0.0
        else:
            nx, ny, jit = 2048 // scale, 1024 // scale, 0.2
        mesh = channel2d(nx, ny, 2.0, 1.0, jitter=jit, seed=SEED0 + cid)
        params = make_params(2, [MELT, GAS_R461], dt=1e-3)
        nc = 2
    elif cid == 2:
        n = 128 // scale
        mesh = kuhn3d(n, n, n, 1.0, 1.0, 1.0)
        params = make_params(3, [MELT, CRYSTAL, H2O], dt=1e-3)
        nc = 3
    elif cid == 4:
        mesh = kuhn3d(192 // scale, 192 // scale, 96 // scale, 2.0, 2.0, 1.0)
        params = make_params(3, [MELT, CRYSTAL, CRYSTAL2, H2O], dt=5e-4)
        nc = 4
    else:
        raise ValueError("config id must be 0..4")
    I, P = fields(mesh, nc, SEED0 + cid)
    return mesh, params, I, P


def mesh_struct(mesh):
    """abi.Mesh view of a mesh dict (keeps references alive in the dict)."""
    import ctypes as C
    arrs = {}
    for k in ("coords", "conn", "node_flag", "side_elem", "side_local", "side_flag"):
        dt = np.float64 if k == "coords" else np.int32
        arrs[k] = np.ascontiguousarray(mesh[k], dtype=dt)
    m = abi.Mesh()
    m.dim = mesh["dim"]
    m.n_nodes = arrs["coords"].shape[0]
    m.n_elems = arrs["conn"].shape[0]
    m.n_sides = arrs["side_elem"].shape[0]
    m.coords = arrs["coords"].ctypes.data_as(C.POINTER(C.c_double))
    for k in ("conn", "node_flag", "side_elem", "side_local", "side_flag"):
        setattr(m, k, arrs[k].ctypes.data_as(C.POINTER(C.c_int)))
    m._keep = arrs
    return m
