"""End-to-end projection steps (LOD transport, pressure RHS, pressure solve,
velocity correction) of the lid-driven cavity on the GPU against the same
steps on the CPU oracle.  The pressure is solved to 1e-13 on both sides so
that the remaining difference is rounding: velocities within 1e-9 relative L2.
"""
import numpy as np
import pytest
import torch

import xf_oracle as xo
import xflow_amd as xf
from xflow_amd import cases
from xflow_amd.drivers import Cavity, velocity_codes

pytestmark = pytest.mark.gpu


def oracle_cavity(n, steps, dt=1e-3, re=100.0):
    h = 1.0 / n
    st = np.full(n + 1, h)
    pcodes = cases.neumann_codes(n, []).astype(np.int32)
    vc, bc = zip(*[velocity_codes(n, a) for a in (1, 2, 3)])
    vel = [b.copy() for b in bc]
    for a in range(3):
        vel[a][vc[a] == cases.PREDEFINED] *= 2.0
    vis = np.full(pcodes.shape, 1.0 / re)
    pg = xo.Grid(st, st, st, pcodes)
    vgs = [xo.Grid(st, st, st, vc[a]) for a in range(3)]
    p = np.zeros(pcodes.shape)
    for _ in range(steps):
        star = []
        for a in range(3):
            tu, tv, tw, tvis, rp = xo.transport_velocity(a + 1, vc[a], *vel, vis=vis, bc=bc[a], h=[st, st, st])
            star.append(xo.lod(vgs[a], vel[a], vel[a], tu, tv, tw, rp, tvis, dt, 0))
        rhs = xo.pressure_rhs(pcodes, star[0], star[1], star[2], st, st, st, dt)
        p, it, _, s = xo.cg_cg(pg, rhs, p, 1e-13, 100000, project_mean=True)
        vel = [xo.velocity_correct(vc[a], a + 1, star[a], p, st, dt, bc=bc[a], out=vel[a]) for a in range(3)]
    return vel, p


def test_cavity_steps_match_oracle():
    ctx = xf.Context(0)
    n, steps = 16, 3
    cav = Cavity(ctx, n=n, tol=1e-13)
    for _ in range(steps):
        info = cav.step()
    vel, p = oracle_cavity(n, steps)
    torch.cuda.synchronize()
    for a in range(3):
        g = cav.vel[a].cpu().numpy()
        d = np.linalg.norm(g - vel[a]) / max(np.linalg.norm(vel[a]), 1e-30)
        assert d <= 1e-9, (a, d)
    assert info["div_max"] < 1e-6


def test_cavity_64_runs_and_stays_divergence_free():
    ctx = xf.Context(0)
    cav = Cavity(ctx, n=64)
    for _ in range(5):
        info = cav.step()
    assert info["div_max"] < 1e-4
    u = cav.vel[0].cpu().numpy()
    assert np.isfinite(u).all() and np.abs(u).max() <= 2.0 + 1e-9


def test_cavity_64_ten_steps_match_oracle():
    """The 64^3 cavity workload itself (Re = 100, dt = 1e-3): ten steps on the
    GPU against the oracle's FP64 restatement of the same steps."""
    ctx = xf.Context(0)
    n, steps = 64, 10
    cav = Cavity(ctx, n=n, tol=1e-13)
    for _ in range(steps):
        info = cav.step()
    vel, p = oracle_cavity(n, steps)
    torch.cuda.synchronize()
    for a in range(3):
        g = cav.vel[a].cpu().numpy()
        d = np.linalg.norm(g - vel[a]) / max(np.linalg.norm(vel[a]), 1e-30)
        assert d <= 1e-9, (a, d)
    assert info["div_max"] < 1e-6
