"""Time-step drivers that exercise the hot path end to end (the reference's
TVPProblem::SolveBySplitting order, XProblem.cpp:4255-6697): per velocity
component an LOD transport step, then the pressure right-hand side, the
pressure solve, the velocity correction and the divergence monitor.

Lid-driven cavity: unit cube with n^3 pressure cells (one-cell ring),
staggered velocities on the pressure array shape (u[i,j,k] on the x-face
between cells i and i+1), no-slip walls, lid u = 1 on the top face (z = 1),
Re = 100 (nu = 0.01), dt = 1e-3, zero initial velocity.

Point types of a velocity component (axis a): faces 0 and n along a are
Defined walls (value 0), the ghost planes 0 and n+1 of the other axes are
PreDefined (the wall value at the mid-point, XOperator.cpp:1057-1068 /
XProblem.cpp:3072-3082), interior faces Actual, edges and the unused plane
n+1 along a Fictive.  The pressure grid is Actual inside a Fictive ring
(Neumann walls by mask).
"""
import numpy as np
import torch

import xflow_amd as xf
from xflow_amd import cases


class StepGraphs:
    """The parts of a time step around the pressure solve, captured once as
    HIP graphs and replayed (no host work or launch gaps per kernel).  With
    graphs on, a whole step runs on one stream of its own, the library's
    context pointed at it too (step()), so that replays, solves and torch
    operations are ordered on that stream alone.  The first call of each part
    runs eagerly (the library allocates its workspaces and checks point types
    then); the second captures it and replays.  The solve itself stays eager:
    its loop is steered from the host."""

    def __init__(self, ctx, enabled):
        self.ctx, self.enabled = ctx, enabled
        self.graphs, self.seen = {}, set()
        self.stream = torch.cuda.Stream() if enabled else None

    def step(self, body):
        if not self.enabled:
            return body()
        torch.cuda.current_stream().synchronize()
        self.ctx.set_stream(self.stream.cuda_stream)
        try:
            with torch.cuda.stream(self.stream):
                return body()
        finally:
            self.ctx.set_stream(None)

    def run(self, name, fn):
        if not self.enabled:
            fn()
            return
        g = self.graphs.get(name)
        if g is None:
            if name not in self.seen:
                self.seen.add(name)
                fn()
                return
            g = torch.cuda.CUDAGraph()
            self.stream.synchronize()
            with torch.cuda.graph(g, stream=self.stream):
                fn()
            self.graphs[name] = g
        g.replay()


def velocity_codes(n, axis, lid_axis=1, lid_value=1.0, top_axis=3):
    """(codes, boundary values) of velocity component `axis` (1..3)."""
    N = n + 2
    idx = np.indices((N, N, N))
    a = axis - 1
    on_wall = (idx[a] == 0) | (idx[a] == n)
    ghost = [(idx[e] == 0) | (idx[e] == n + 1) for e in range(3) if e != a]
    count = on_wall.astype(int) + sum(g.astype(int) for g in ghost)
    codes = np.full((N, N, N), cases.ACTUAL, dtype=np.int32)
    codes[count == 1] = cases.PREDEFINED
    codes[(count == 1) & on_wall] = cases.DEFINED
    codes[count >= 2] = cases.FICTIVE
    codes[idx[a] == n + 1] = cases.FICTIVE
    bc = np.zeros((N, N, N))
    if axis == lid_axis:
        t = top_axis - 1
        bc[(idx[t] == n + 1) & (codes == cases.PREDEFINED)] = lid_value
    return codes, bc


class Cavity:
    """GPU lid-driven cavity on libxflow_hip.so."""

    def __init__(self, ctx, n=64, re=100.0, dt=1e-3, tol=1e-8, solver=None, graphs=False):
        self.ctx, self.n, self.dt, self.tol = ctx, n, dt, tol
        self.solver = xf.XF_CG if solver is None else solver
        self.graphs = StepGraphs(ctx, graphs)
        h = 1.0 / n
        self.h = h
        steps = np.full(n + 1, h)
        self.steps = steps
        self.pcodes = cases.neumann_codes(n, []).astype(np.int32)
        self.gp = xf.Grid(ctx, steps, steps, steps, self.pcodes)
        self.vgrids, self.vcodes, self.bc = [], [], []
        for axis in (1, 2, 3):
            codes, bc = velocity_codes(n, axis)
            self.vcodes.append(codes)
            self.bc.append(torch.from_numpy(bc).cuda())
            self.vgrids.append(xf.Grid(ctx, steps, steps, steps, codes))
        shape = self.pcodes.shape
        z = lambda: torch.zeros(shape, dtype=torch.float64, device="cuda")
        # fields: current velocity, predictor, previous, transport velocities, pressure
        self.vel = [self.bc[a].clone() for a in range(3)]
        for a in range(3):  # PreDefined ghosts hold 2 bc - inner = 2 bc at rest
            self.vel[a][torch.from_numpy(self.vcodes[a] == cases.PREDEFINED).cuda()] *= 2.0
        self.star = [z() for _ in range(3)]
        # per component: transport velocities, the viscosity's two-node mean and
        # the boundary right-hand side, built as the reference builds them from
        # the viscosity at the pressure nodes
        self.trs = [[z() for _ in range(3)] for _ in range(3)]
        self.vis = torch.full(shape, 1.0 / re, dtype=torch.float64, device="cuda")
        self.tviss = [z() for _ in range(3)]
        self.rps = [z() for _ in range(3)]
        self.p = z()
        self.p0 = z()
        self.rhs = z()
        self.last = None

    def _predict(self):
        # velocity predictor: LOD transport per component (XProblem.cpp:5017-5446),
        # the three components' LOD steps in one set of launches
        for a in range(3):
            self.vgrids[a].transport_terms(a + 1, self.vel, self.trs[a], vis=self.vis, bc=self.bc[a],
                                           tvis=self.tviss[a], rp=self.rps[a])
            self.star[a].copy_(self.vel[a])
        xf.lod_transport_batch(self.ctx, self.vgrids, self.star, self.vel, *zip(*self.trs), self.rps, self.tviss,
                               self.dt, 0)
        # pressure right-hand side (XProblem.cpp:5474-5513)
        self.gp.pressure_rhs(self.star[0], self.star[1], self.star[2], self.steps, self.steps, self.steps, self.dt,
                             self.rhs)

    def _correct(self):
        # velocity correction (XProblem.cpp:5598-5778)
        for a in range(3):
            self.vgrids[a].velocity_correct(a + 1, self.star[a], self.p, self.steps, self.dt, self.vel[a],
                                            bc=self.bc[a])

    def step(self):
        return self.graphs.step(self._step)

    def _step(self):
        self.graphs.run("predict", self._predict)
        o = xf.solve_opts(self.solver, tol=self.tol, tol_mode=xf.XF_TOL_REL, project_mean=1)
        st = self.gp.solve(self.rhs, self.p0, self.p, o)
        if st.status != 0:
            raise xf.XfError(st.status, "pressure solve failed")
        self.p0.copy_(self.p)
        # the velocity correction's boundary-value checks are read back once,
        # after the divergence monitor's wait
        self.ctx.set_deferred_errors(True)
        try:
            self.graphs.run("correct", self._correct)
            div, _ = self.gp.divergence_max(self.vel[0], self.vel[1], self.vel[2], self.steps, self.steps,
                                            self.steps)
        finally:
            self.ctx.set_deferred_errors(False)
        self.ctx.check_errors()
        self.last = {"pressure_iters": st.iters, "pressure_ms": st.ms, "div_max": div,
                     "resid_true": st.resid_true}
        return self.last


def channel_velocity_codes(nx, ny, nz, axis):
    """Velocity point types of the IB channel: inflow/outflow planes Defined
    (u = 1 for the streamwise component), no-slip walls y, z; the streamwise
    ghost planes of v and w PreDefined 0."""
    shape = (nx + 2, ny + 2, nz + 2)
    E = (nx, ny, nz)
    idx = np.indices(shape)
    a = axis - 1
    on_wall = (idx[a] == 0) | (idx[a] == E[a])
    ghost = [(idx[e] == 0) | (idx[e] == E[e] + 1) for e in range(3) if e != a]
    count = on_wall.astype(int) + sum(g.astype(int) for g in ghost)
    codes = np.full(shape, cases.ACTUAL, dtype=np.int32)
    codes[count == 1] = cases.PREDEFINED
    codes[(count == 1) & on_wall] = cases.DEFINED
    codes[count >= 2] = cases.FICTIVE
    codes[idx[a] == E[a] + 1] = cases.FICTIVE
    bc = np.zeros(shape)
    if axis == 1:
        bc[codes == cases.DEFINED] = 1.0  # plug inflow and outflow, u = 1
    return codes, bc


def channel_scalar_codes(nx, ny, nz):
    """Cell-centred scalar: inflow ghost PreDefined (c = 0), outflow and wall
    ghosts PreNormal (zero flux: the reference's concentration boundary,
    XProblem.cpp:5925-5929), edges Fictive."""
    shape = (nx + 2, ny + 2, nz + 2)
    E = (nx, ny, nz)
    idx = np.indices(shape)
    ghost = [(idx[e] == 0) | (idx[e] == E[e] + 1) for e in range(3)]
    count = sum(g.astype(int) for g in ghost)
    codes = np.full(shape, cases.ACTUAL, dtype=np.int32)
    codes[count == 1] = cases.PRENORMAL
    codes[(count == 1) & (idx[0] == 0)] = cases.PREDEFINED
    codes[count >= 2] = cases.FICTIVE
    return codes


class IBChannel:
    """Flow past a sphere in a channel with direct-forcing immersed boundary
    and a passive scalar (the 256x128x128 IB workload at any size).

    Channel [0, 2H] x [0, H] x [0, H] with H = 1, n_y = n_z = ny cells; sphere
    D = 0.25 H at (0.5, 0.5, 0.5) as an icosphere (subdivision 4: 2562
    vertices); inflow u = 1, Re = 200 on D (nu = U D / Re); scalar Pe = 500
    (kappa = U D / Pe), Gaussian blob (sigma 0.05) at (0.25, 0.5, 0.5);
    dt = 0.5 h / U (CFL 0.5).  Each step: LOD predictor per component, direct
    forcing F = (0 - U(X)) / dt at the vertices spread with the vertex
    volumes (area x h), pressure projection, LOD transport of the scalar."""

    def __init__(self, ctx, ny=128, subdivisions=4, tol=1e-8, solver=None, precond=None, graphs=False):
        self.ctx = ctx
        self.graphs = StepGraphs(ctx, graphs)
        nx, nz = 2 * ny, ny
        self.E = (nx, ny, nz)
        h = 1.0 / ny
        self.h = h
        self.dt = 0.5 * h / 1.0
        D = 0.25
        self.nu, self.kap = 1.0 * D / 200.0, 1.0 * D / 500.0
        self.tol = tol
        self.solver = xf.XF_CG if solver is None else solver
        self.precond = precond
        self.verify = False  # report each solve's true residual (one more FP64 operator pass)
        self.steps = [np.full(e + 1, h) for e in self.E]
        shape = (nx + 2, ny + 2, nz + 2)
        self.pcodes = np.zeros(shape, dtype=np.int32)
        self.pcodes[1:-1, 1:-1, 1:-1] = cases.ACTUAL
        self.gp = xf.Grid(ctx, *self.steps, self.pcodes)
        self.vgrids, self.bc, self.vcodes = [], [], []
        for axis in (1, 2, 3):
            codes, bc = channel_velocity_codes(nx, ny, nz, axis)
            self.vcodes.append(codes)
            self.vgrids.append(xf.Grid(ctx, *self.steps, codes))
            self.bc.append(torch.from_numpy(bc).cuda())
        self.scodes = channel_scalar_codes(nx, ny, nz)
        self.gs = xf.Grid(ctx, *self.steps, self.scodes)
        V, F = cases.icosphere(subdivisions, D / 2, (0.5, 0.5, 0.5))
        self.nodes = torch.from_numpy(np.ascontiguousarray(V)).cuda()
        self.vol = torch.from_numpy(cases.vertex_areas(V, F) * h).cuda()
        z = lambda: torch.zeros(shape, dtype=torch.float64, device="cuda")
        # initial state: plug flow u = 1 in the fluid, ghosts consistent with the walls
        self.vel = [z() for _ in range(3)]
        u0 = np.zeros(shape)
        u0[self.vcodes[0] == cases.ACTUAL] = 1.0
        u0[self.vcodes[0] == cases.DEFINED] = 1.0
        pre = self.vcodes[0] == cases.PREDEFINED
        u0[pre] = -1.0  # mirror of the inner u = 1 about the no-slip wall value 0
        self.vel[0] = torch.from_numpy(u0).cuda()
        self.star = [z() for _ in range(3)]
        self.trs = [[z() for _ in range(3)] for _ in range(3)]  # per component (the last: the scalar's too)
        self.force = z()
        self.nu_f = torch.full(shape, self.nu, dtype=torch.float64, device="cuda")
        self.kap_f = torch.full(shape, self.kap, dtype=torch.float64, device="cuda")
        self.tviss = [z() for _ in range(3)]
        self.rps = [z() for _ in range(3)]
        X = (np.arange(nx + 2) - 0.5) * h
        Y = (np.arange(ny + 2) - 0.5) * h
        Zc = (np.arange(nz + 2) - 0.5) * h
        XX, YY, ZZ = np.meshgrid(X, Y, Zc, indexing="ij")
        c0 = np.exp(-((XX - 0.25) ** 2 + (YY - 0.5) ** 2 + (ZZ - 0.5) ** 2) / (2 * 0.05 ** 2))
        c0[self.scodes != cases.ACTUAL] = 0.0
        self.c = torch.from_numpy(c0).cuda()
        self.c_new = z()
        self.zero = z()
        self.p, self.p0, self.rhs = z(), z(), z()
        self.U = torch.zeros(self.nodes.shape[0], dtype=torch.float64, device="cuda")
        self.last = None

    def _predict(self):
        dt = self.dt
        for a in range(3):  # velocity predictor; the components' LOD steps in one set of launches
            self.vgrids[a].transport_terms(a + 1, self.vel, self.trs[a], vis=self.nu_f, bc=self.bc[a],
                                           tvis=self.tviss[a], rp=self.rps[a])
            self.star[a].copy_(self.vel[a])
        xf.lod_transport_batch(self.ctx, self.vgrids, self.star, self.vel, *zip(*self.trs), self.rps, self.tviss,
                               dt, 0)
        for a in range(3):  # direct forcing at the sphere's vertices
            g = self.vgrids[a]
            g.ib_interp(a + 1, self.nodes, self.star[a], self.U)
            F = (-self.U / dt) * self.vol
            g.ib_spread(a + 1, self.nodes, F, self.force)
            g.lincomb(1.0, self.star[a], dt, self.force, self.star[a])
        self.gp.pressure_rhs(self.star[0], self.star[1], self.star[2], *self.steps, dt, self.rhs)

    def _correct(self):
        dt = self.dt
        for a in range(3):
            self.vgrids[a].velocity_correct(a + 1, self.star[a], self.p, self.steps[a], dt, self.vel[a], bc=self.bc[a])
        # passive scalar, upwind LOD with cell-centred velocities
        tr, rp = self.trs[2], self.rps[2]
        self.gs.transport_terms(0, self.vel, tr, bc=self.zero, rp=rp)
        self.c_new.copy_(self.c)
        self.gs.lod_transport(self.c_new, self.c, tr[0], tr[1], tr[2], rp, self.kap_f, dt, 1)
        self.c.copy_(self.c_new)

    def step(self):
        return self.graphs.step(self._step)

    def _step(self):
        self.graphs.run("predict", self._predict)
        o = xf.solve_opts(self.solver, tol=self.tol, tol_mode=xf.XF_TOL_REL, project_mean=1)
        o.verify = 1 if self.verify else 0
        if self.precond is not None:
            o.precond = self.precond
        st = self.gp.solve(self.rhs, self.p0, self.p, o)
        if st.status != 0:
            raise xf.XfError(st.status, "pressure solve failed")
        self.p0.copy_(self.p)
        # the velocity correction's boundary-value checks are read back once,
        # after the divergence monitor's wait
        self.ctx.set_deferred_errors(True)
        try:
            self.graphs.run("correct", self._correct)
            div, _ = self.gp.divergence_max(self.vel[0], self.vel[1], self.vel[2], *self.steps)
        finally:
            self.ctx.set_deferred_errors(False)
        self.ctx.check_errors()
        self.last = {"pressure_iters": st.iters, "inner_iters": st.inner_iters, "pressure_ms": st.ms,
                     "div_max": div, "resid_true": st.resid_true}
        return self.last
