"""The multigrid start of the pressure solve (rbgs_mg_cycles; no reference
counterpart): W-cycles of a truncated aggregation multigrid correct the start
before the red-black iteration runs to the tolerance.

CPU, against what the mathematics fixes:
- each level's operator is the Galerkin product P^T A P of the free cells'
  operator with piecewise-constant 2x2x2 aggregation (scipy, in float64);
- the cycles, repeated, converge to the exact solution of the linear system
  over the free cells (scipy's direct solve) at a rate well below one per
  cycle, and hold the cells at the pressure clamp and the non-fluid cells;
- on the 64^3 pour, the start cuts the iterations to the 1e-4 tolerance by
  more than 10x.
GPU: the device start (fine level over the active bricks, whole-grid middle
levels, the small levels inside one workgroup, with their rows in registers,
with the level above them in the same workgroup, or from the operation list)
equals the oracle's bit for bit: iteration counts
and the whole state, from zero, previous-step and external starts, on the
persistent and the launch-per-chunk solve."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import coffee_sim_amd as cs
from helpers import compare_state

FP = 2.0 ** 18
CLAMP = 8191.0  # pressure_clamp: FP_VALUE_LIMIT - 1 (state.rs:15-18, mod.rs:592)


def _state(orc, scale=4, frames=3, seed=6):
    """An oracle stopped after classify of a pour substep, its previous-step
    pressure on the fluid cells as the current pressure."""
    w = cs.SeededWorkload("v60_working", seed=seed, scale=scale)
    s = w.settings.copy()
    s.warm_start = cs.abi.CS_WARM_PREVIOUS
    o = orc.OracleSim(s, w.scene)
    for _ in range(frames):
        o.step_frame(w.frame_dt)
    prev, _ = o.read_pressure()
    o.debug_step_until(w.frame_dt, "classify")
    kind = o.debug_read_grid_lanes()[3]
    fluid = (kind >= 1) & (kind <= 3)
    p0 = np.where(fluid, prev, 0).astype(np.int32)
    o.debug_set_pressure(p0)
    return w, s, o, p0, fluid


def _system(o, p0, fluid, clamp):
    """Sparse A (free x free), b, the free mask and neighbour index table."""
    W, cnt, c = o.read_system()
    X, Y, Z = o.dims
    nc = X * Y * Z
    pf = p0.astype(np.float64) / FP
    free = fluid & (cnt > 0) & (np.abs(pf) < clamp)
    idx = np.arange(nc)
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    rows, cols, vals = [], [], []
    for k, (dx, dy, dz) in enumerate([(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]):
        nx, ny, nz = x + dx, y + dy, z + dz
        ok = (nx >= 0) & (nx < X) & (ny >= 0) & (ny < Y) & (nz >= 0) & (nz < Z)
        j = np.where(ok, (nz * Y + ny) * X + nx, 0)
        m = ok & free & free[j] & (W[:, k] > 0)
        rows.append(idx[m]); cols.append(j[m]); vals.append(W[m, k].astype(np.float64))
    off = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nc, nc))
    A = (sp.diags(np.where(free, cnt, 0.0).astype(np.float64)) - off).tocsr()
    return A, -c.astype(np.float64), free, cnt, W


def _aggregate(dims):
    X, Y, Z = dims
    Xc, Yc, Zc = (X + 1) // 2, (Y + 1) // 2, (Z + 1) // 2
    idx = np.arange(X * Y * Z)
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    parent = ((z // 2) * Yc + y // 2) * Xc + x // 2
    return sp.csr_matrix((np.ones(len(idx)), (idx, parent)), shape=(len(idx), Xc * Yc * Zc)), (Xc, Yc, Zc)


def _faces(Ac, dims):
    """Diagonal and +x/+y/+z couplings (negated off-diagonals) of a dense-grid operator."""
    X, Y, Z = dims
    n = X * Y * Z
    i = np.arange(n)
    x, y, z = i % X, (i // X) % Y, i // (X * Y)
    D = Ac.diagonal()
    out = [D]
    for (dx, dy, dz), lim in (((1, 0, 0), x + 1 < X), ((0, 1, 0), y + 1 < Y), ((0, 0, 1), z + 1 < Z)):
        j = i + dx + dy * X + dz * X * Y
        w = np.zeros(n)
        w[lim] = -np.asarray(Ac[i[lim], j[lim]]).ravel()
        out.append(w)
    return [a.reshape(Z, Y, X) for a in out]


def test_oracle_levels_are_galerkin_products(orc):
    w, s, o, p0, fluid = _state(orc)
    A, b, free, cnt, W = _system(o, p0, fluid, CLAMP)
    dims = o.dims
    Ac = A
    for level in range(0, 4):
        got = o.debug_mg_operator(w.frame_dt, level)
        if got is None:
            break
        want = _faces(Ac, dims)
        for name, g, e in zip(("D", "Wx", "Wy", "Wz"), got, want):
            e = np.maximum(e, 0.0) if name == "D" else e
            scale = max(1.0, float(np.abs(e).max()))
            assert np.abs(g.astype(np.float64) - e).max() <= 2e-6 * scale * (8 ** level), (level, name)
        assert (got[0] > 0).sum() > 0
        P, dims = _aggregate(dims)
        Ac = (P.T @ Ac @ P).tocsr()
    assert level >= 2


def test_oracle_cycles_converge_to_the_linear_solution(orc):
    w, s, o, p0, fluid = _state(orc)
    clamp = CLAMP
    A, b, free, cnt, W = _system(o, p0, fluid, clamp)
    pf = p0.astype(np.float64) / FP
    assert (np.abs(pf[fluid]) < clamp).all()  # no cell starts at the clamp in this state
    f = np.nonzero(free)[0]
    x = pf.copy()
    x[f] = spla.spsolve(A[f][:, f].tocsc(), b[f])
    assert np.abs(x).max() < clamp

    def rel(p):
        r = (b - A @ p)[f]
        return np.linalg.norm(r) / np.linalg.norm(b[f])

    res = []
    for cycles in (1, 2, 4, 12):
        o.debug_set_pressure(p0)
        o.set_mg_cycles(cycles)
        o.debug_run_pass(w.frame_dt, "mg")
        p = o.debug_read_grid_lanes()[0].astype(np.float64) / FP
        assert (p[~fluid] == 0).all()
        res.append(rel(p))
    r0 = rel(pf)
    assert res[0] < 0.5 * r0
    assert res[2] < 0.05 * res[0]  # three more cycles: well below 0.5 per cycle
    assert res[3] < 1e-4
    assert np.abs(p - x).max() <= 1e-3 * np.abs(x).max()


def test_oracle_cycles_hold_active_clamped_cells(orc):
    """Within a cycle, a cell at the clamp whose unclamped update (sum - c) /
    count stays beyond the bound is held; one whose update lies back inside
    takes part (it may move).  (The next cycle judges again from the moved
    neighbours.)  Non-fluid cells stay 0."""
    w, s, o, p0, fluid = _state(orc)
    W, cnt, c = o.read_system()
    X, Y, Z = o.dims
    rng = np.random.default_rng(3)
    lim = int(CLAMP * FP)
    p = p0.copy()
    nb = [-1, 1, -X, X, -X * Y, X * Y]
    inner = [i for i in np.nonzero(fluid & (cnt > 0))[0]
             if X < i % (X * Y) < X * (Y - 1) and 0 < i % X < X - 1 and X * Y <= i < X * Y * (Z - 1)]
    centres = rng.choice(np.array(inner), 60, replace=False)
    for k, i in enumerate(centres):  # the cell and its face neighbours at one bound
        v = lim if k % 2 == 0 else -lim
        for d in [0] + nb:
            if fluid[i + d]:
                p[i + d] = v
    pf = p.astype(np.float64) / FP
    held, free = [], []
    for i in centres:
        ssum = sum(float(W[i, k]) * pf[i + nb[k]] for k in range(6) if W[i, k] > 0 and fluid[i + nb[k]])
        v = (ssum - float(c[i])) / float(cnt[i])
        at = np.sign(pf[i])
        if at * v > CLAMP * (1 + 1e-6):
            held.append(i)
        elif at * v < CLAMP * (1 - 1e-6):
            free.append(i)
    assert held and free, (len(held), len(free))
    o.debug_set_pressure(p)
    o.set_mg_cycles(1)
    o.debug_run_pass(w.frame_dt, "mg")
    got = o.debug_read_grid_lanes()[0]
    assert (got[held] == p[held]).all()
    assert (got[free] != p[free]).any()
    assert (got[~fluid] == 0).all()


def test_oracle_start_cuts_iterations_on_the_pour(orc):
    w = cs.SeededWorkload("v60_working", seed=1, scale=2)
    its = {}
    for cycles in (0, 3):
        s = w.settings.copy()
        s.warm_start = cs.abi.CS_WARM_PREVIOUS
        s.rbgs_mg_cycles = cycles
        o = orc.OracleSim(s, w.scene)
        its[cycles] = []
        for _ in range(4):
            o.step_frame(w.frame_dt)
            st = o.solve_stats()
            assert st.converged
            its[cycles].append(int(st.iterations))
        o.close()
    assert sum(its[0][1:]) > 10 * max(sum(its[3][1:]), 1), its


# ---- GPU ---------------------------------------------------------------------------------

def _pair(ctx, orc, s, scene):
    return cs.MpmSim3D(ctx, s, scene), orc.OracleSim(s, scene)


@pytest.mark.gpu
@pytest.mark.parametrize("visit", ["rows3", "rows", "oplist", "xfastest"])
@pytest.mark.parametrize("layout", ["persistent", "launches"])
@pytest.mark.parametrize("scale", [4, 2])
def test_mg_start_matches_oracle(ctx, orc, monkeypatch, layout, scale, visit):
    """visit: the small levels' visit with its rows in registers, joined by the
    level above them where that is not the fine level (the default; on the
    32^3 grid the same as "rows"), the two-level register visit
    (CS_MG_ROWS=1), or the operation-list visit (CS_MG_ROWS=0).  The fine
    level's arrays are brick-major, or (xfastest, CS_MG_BRICKS=0) laid out
    like the grid."""
    if layout == "launches":
        monkeypatch.setenv("CS_RBGS_PERSIST", "0")
    if visit == "xfastest":
        monkeypatch.setenv("CS_MG_BRICKS", "0")
    monkeypatch.setenv("CS_MG_ROWS", {"rows3": "2", "rows": "1", "oplist": "0", "xfastest": "2"}[visit])
    w = cs.SeededWorkload("v60_working", seed=6, scale=scale)
    nc = int(np.prod(list(w.settings.grid_dims)))
    starts = (cs.abi.CS_WARM_ZERO, cs.abi.CS_WARM_PREVIOUS, cs.abi.CS_WARM_EXTERNAL) if scale == 4 else \
        (cs.abi.CS_WARM_PREVIOUS,)
    for start in starts:
        s = w.settings.copy()
        s.warm_start = start
        s.rbgs_mg_cycles = 3
        s.warm_fallback_residual = 1e30
        field = None
        if start == cs.abi.CS_WARM_EXTERNAL:
            rng = np.random.default_rng(7)
            field = rng.normal(0.0, 2.0, nc).astype(np.float32)
        sim, ref = _pair(ctx, orc, s, w.scene)
        for k in range(3):
            if field is not None:
                sim.set_warm_start(field)
                ref.set_warm_start(field)
            sim.step_frame(w.frame_dt)
            ref.step_frame(w.frame_dt)
            sg, so = sim.solve_stats(), ref.solve_stats()
            assert sg.iterations == so.iterations, (layout, visit, scale, start, k, sg.as_dict(), so.as_dict())
            assert sg.converged == so.converged == 1
            compare_state(sim, ref, f"mg start {layout} {visit} scale {scale} start {start} substep {k}")
        sim.close()
        ref.close()


@pytest.mark.gpu
def test_mg_start_cuts_device_iterations(ctx):
    """On the 64^3 pour the device start cuts the iterations to 1e-4 more than
    10x against the previous-step start alone, and the kernel timing reports it."""
    w = cs.SeededWorkload("v60_working", seed=1, scale=2)
    its = {}
    for cycles in (0, 3):
        s = w.settings.copy()
        s.warm_start = cs.abi.CS_WARM_PREVIOUS
        s.rbgs_mg_cycles = cycles
        sim = cs.MpmSim3D(ctx, s, w.scene)
        sim.enable_timing(True)
        sim.reset_kernel_timing()
        its[cycles] = []
        for _ in range(6):
            sim.step_frame(w.frame_dt)
            st = sim.solve_stats()
            assert st.converged
            its[cycles].append(int(st.iterations))
        kt = sim.kernel_timing()
        assert (kt.mg_starts > 0) == (cycles > 0)
        sim.close()
    assert sum(its[0][1:]) > 10 * max(sum(its[3][1:]), 1), its


@pytest.mark.gpu
def test_mg_cycles_are_validated(ctx):
    w = cs.SeededWorkload("v60_working", seed=1, scale=4)
    s = w.settings.copy()
    s.rbgs_mg_cycles = 2
    s.rbgs_mode = cs.abi.CS_RBGS_FIXED_PAIRS
    with pytest.raises(cs.CoffeeSimError):
        cs.MpmSim3D(ctx, s, w.scene)
    s = w.settings.copy()
    s.rbgs_mg_cycles = 2
    s.rbgs_rescale_after = 32
    with pytest.raises(cs.CoffeeSimError):
        cs.MpmSim3D(ctx, s, w.scene)
