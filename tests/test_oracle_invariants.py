"""The oracle's passes against what the mathematics fixes (closed forms,
conservation, a direct solve of the same linear system).  Each test is chosen
so a plausible slip (a dropped term, a wrong sign or index, a transposed
operand) fails it."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from scenes import DX, SceneArrays, bed_field, block_particles, box_settings, cell_index

FP = 2.0 ** 18
G = None


def bspline_weights(fx):
    return np.array([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])


def one_particle(orc, pos, vel, C=None, J=1.0, rbgs_pairs=0):
    s = box_settings(n=16, rbgs_pairs=rbgs_pairs)
    p = np.array([[pos[0], pos[1], pos[2], J, vel[0], vel[1], vel[2], 0.25]], np.float32)
    a = np.zeros((1, 12), np.float32)
    if C is not None:
        a[0, 0:3], a[0, 4:7], a[0, 8:11] = C[0], C[1], C[2]
    sc = SceneArrays(p, a)
    return s, orc.OracleSim(s, sc.scene), sc


def test_p2g_partition_of_unity_and_first_moment(orc):
    rng = np.random.default_rng(3)
    pos = (rng.random(3) - 0.5) * 0.6
    vel = np.array([1.0, -2.0, 0.5])
    Cm = rng.normal(size=(3, 3)) * 4.0
    s, o, _ = one_particle(orc, pos, vel, C=Cm, J=1.2)
    o.debug_step_until(1e-3, "p2g")
    lanes = o.read_p2g_lanes().astype(np.float64) / FP
    m = 0.25
    tol = 27.0 / FP + 1e-6
    # sum_i w_i = 1: mass, rest and current volume are conserved up to the
    # 27 truncations toward zero
    assert abs(lanes[0].sum() - m) <= tol
    p_vol = DX ** 3 * 0.125
    assert abs(lanes[4].sum() - p_vol) <= tol
    assert abs(lanes[5].sum() - 1.2 * p_vol) <= tol
    # sum_i w_i (x_i - x_p) = 0: the affine term carries no net momentum
    for c in range(3):
        assert abs(lanes[1 + c].sum() - m * vel[c]) <= 27.0 / FP + 1e-5 * (abs(m * vel[c]) + 1.0)
    # exactly the 27 stencil nodes are touched
    assert np.count_nonzero(lanes[0]) == 27


def test_p2g_node_weights_and_reference_affine_convention(orc):
    """Per-node deposits against float64 B-spline weights.  The reference
    deposits mom.x += w * m * (col0 . dpos) (shader.rs:1018-1022, affine
    columns as written by G2P at shader.rs:1750-1752); this pins that
    orientation of C."""
    pos = np.array([0.1234, -0.0871, 0.0412])
    vel = np.array([0.3, 0.2, -0.1])
    Cm = np.zeros((3, 3))
    Cm[0] = [0.0, 5.0, 0.0]  # col0 = (0, 5, 0)
    s, o, _ = one_particle(orc, pos, vel, C=Cm)
    o.debug_step_until(1e-3, "p2g")
    lanes = o.read_p2g_lanes()
    n = s.grid_dims[0]
    g = (pos + 0.5 * n * DX) / DX
    base = np.floor(g - 0.5).astype(int)
    fx = g - base
    w = [bspline_weights(fx[c]) for c in range(3)]
    m = 0.25
    for i in range(3):
        for j in range(3):
            for k in range(3):
                ww = w[0][i] * w[1][j] * w[2][k]
                dpos = (np.array([i, j, k]) - fx) * DX
                ci = cell_index(s, base[0] + i, base[1] + j, base[2] + k)
                assert abs(lanes[0, ci] - np.trunc(ww * m * FP)) <= 1
                mom = ww * (m * vel + m * (Cm @ dpos))  # row c of Cm = column c of C
                for c in range(3):
                    assert abs(lanes[1 + c, ci] - np.trunc(mom[c] * FP)) <= 2, (i, j, k, c)


def test_grid_update_closed_form(orc):
    pos = np.array([0.05, 0.11, -0.07])
    vel = np.array([0.7, -0.4, 0.2])
    s, o, _ = one_particle(orc, pos, vel)
    dt = 1e-3
    o.debug_step_until(dt, "grid")
    gv = o.read_grid_vel()
    lanes = o.read_p2g_lanes()
    touched = lanes[0] > 0
    moved = lanes[0] / FP >= 1e-6
    assert 20 <= moved.sum() <= touched.sum() <= 27
    # nodes below the 1e-6 mass floor keep the cleared grid_vel (shader.rs:1064)
    assert np.all(gv[touched & ~moved] == 0.0)
    want = vel + np.array([0.0, s.gravity * dt, 0.0])
    # v = (trunc(w m v fp) / fp) / (trunc(w m fp) / fp): each truncation moves
    # the ratio by at most (1 + |v|) / (mass lane)
    bound = (1.0 + np.abs(vel).max()) / lanes[0, moved] + 1e-6
    err = np.abs(gv[moved, :3] - want).max(axis=1)
    assert np.all(err <= bound), (err, bound)
    assert np.allclose(gv[moved, 3], lanes[0, moved] / FP, rtol=1e-7)


def test_free_fall_without_pressure(orc):
    """0 RBGS pairs: P2G -> grid -> G2P reproduces ballistic motion."""
    pos = np.array([0.02, 0.3, -0.04])
    vel = np.array([0.9, 0.0, -0.3])
    s, o, _ = one_particle(orc, pos, vel, rbgs_pairs=0)
    dt = 1e-3
    o.step_frame(dt)
    p = o.read_particles()[0].astype(np.float64)
    want_v = vel + np.array([0.0, s.gravity * dt, 0.0])
    assert np.abs(p[4:7] - want_v).max() < 2e-3 * abs(s.gravity * dt)
    assert np.abs(p[0:3] - (pos + p[4:7] * dt)).max() < 1e-6
    assert abs(p[3] - 1.0) < 1e-3


def _pool(orc, tol=1e-5, **kw):
    s = box_settings(n=16, cup=True, tolerance=tol, **kw)
    n = s.grid_dims[0]
    r_cells = 0.42 * n - 1.5

    def inside(ix, iy, iz):
        return (ix + 0.5 - n / 2) ** 2 + (iz + 0.5 - n / 2) ** 2 <= r_cells ** 2

    parts = block_particles(s, (0, 3, 0), (n, 9, n), inside=inside)
    sc = SceneArrays(parts)
    return s, orc.OracleSim(s, sc.scene), sc


@pytest.fixture(scope="module")
def pool_after_solve(orc):
    s, o, sc = _pool(orc)
    o.debug_step_until(1e-3, "rbgs")
    return s, o, sc


def test_rbgs_converges_to_direct_solve(pool_after_solve):
    s, o, _ = pool_after_solve
    st = o.solve_stats()
    assert st.converged == 1 and st.final_residual <= 1e-5
    w, count, c = o.read_system()
    p_fp, kind = o.read_pressure()
    fluid = np.nonzero((kind >= 1) & (kind <= 3) & (count > 0))[0]
    assert fluid.size > 200
    pos = {int(i): k for k, i in enumerate(fluid)}
    n = s.grid_dims[0]
    off = [-1, 1, -n, n, -n * n, n * n]
    rows, cols, vals = [], [], []
    for k, i in enumerate(fluid):
        rows.append(k); cols.append(k); vals.append(float(count[i]))
        for d in range(6):
            j = int(i) + off[d]
            if w[i, d] > 0 and j in pos:
                rows.append(k); cols.append(pos[j]); vals.append(-float(w[i, d]))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(fluid.size, fluid.size))
    b = -c[fluid].astype(np.float64)
    p_star = spla.spsolve(A.tocsc(), b)
    p = p_fp[fluid] / FP
    scale = np.abs(p_star).max()
    assert scale > 1.0
    assert np.abs(p - p_star).max() <= 1e-3 * scale
    # the reported residual is ||b - A p|| / ||b|| of this very system
    r = np.linalg.norm(b - A @ (p_fp[fluid] / FP)) / np.linalg.norm(b)
    assert abs(r - st.final_residual) <= 1e-6 * max(r, 1e-12) + 1e-12


def test_pressure_weights_of_interior_cells(pool_after_solve):
    s, o, _ = pool_after_solve
    w, count, c = o.read_system()
    _, kind = o.read_pressure()
    interior = np.nonzero(kind == 2)[0]
    assert interior.size > 50
    # all six neighbours of an interior cell are fluid or solid; fluid faces
    # weigh min(1, fill) and solid faces drop out (Neumann)
    assert np.all(count[interior] <= 6.0)
    full = interior[np.all(w[interior] == 1.0, axis=1)]
    assert full.size > 0 and np.all(count[full] == 6.0)


def test_projection_opposes_gravity_in_a_resting_pool(orc):
    """A resting pool: grid_update adds g*dt to v_y.  The converged pressure
    rises monotonically with depth in every interior column and its gradient
    (shader.rs:1517-1523) pushes the pool up: the mass-weighted v_y after the
    projection is well above g*dt (a flipped gradient would push it below)."""
    s, o, _ = _pool(orc)
    dt = 1e-3
    o.debug_step_until(dt, "classify")
    before = o.read_grid_vel().copy()
    o.debug_step_until(dt, "project")
    gv = o.read_grid_vel()
    p, kind = o.read_pressure()
    n = s.grid_dims[0]
    gdt = s.gravity * dt
    kk = kind.reshape(n, n, n)
    pp = p.reshape(n, n, n)
    cols = 0
    for iz in range(n):
        for ix in range(n):
            ys = np.nonzero(kk[iz, :, ix] == 2)[0]
            if ys.size >= 4:
                cols += 1
                assert np.all(np.diff(pp[iz, ys, ix]) < 0)
    assert cols > 20
    fluid = (kind >= 1) & (kind <= 3)
    m = before[fluid, 3]
    vy_before = np.sum(m * before[fluid, 1]) / m.sum()
    vy_after = np.sum(m * gv[fluid, 1]) / m.sum()
    assert vy_before < 0.95 * gdt  # everything falls at g*dt before the projection
    assert vy_after > 0.5 * gdt, (vy_after, gdt)


def test_grid_darcy_damping_closed_form(orc):
    """Bed cells with zero pressure gradient: v <- v / (1 + nu dt / k)
    (shader.rs:1524-1533) with v_bed = 0."""
    s = box_settings(n=16, rbgs_pairs=0, bulk=0.0)
    s.has_bed_field = 1
    parts = block_particles(s, (4, 4, 4), (12, 10, 12), vel=(0.5, -1.0, 0.25))
    lookup, recs = bed_field(s, (4, 4, 4), (12, 10, 12), porosity=0.4, permeability=2e-9)
    sc = SceneArrays(parts, bed_lookup=lookup, bed_records=recs)
    o = orc.OracleSim(s, sc.scene)
    dt = 1e-3
    o.debug_step_until(dt, "classify")
    before = o.read_grid_vel().copy()
    o.debug_step_until(dt, "project")
    after = o.read_grid_vel()
    _, kind = o.read_pressure()
    bed = np.nonzero(kind == 3)[0]
    assert bed.size > 100
    # away from walls the boundary pass leaves v alone: compare the ratio
    damping = 1.0 / (1.0 + 1e-6 / 2e-9 * dt)
    ok = np.abs(before[bed, 0]) > 1e-3
    ratio = after[bed, 0][ok] / before[bed, 0][ok]
    assert np.allclose(ratio, damping, rtol=1e-5)


def test_viscosity_keeps_a_uniform_field(orc):
    s = box_settings(n=16, rbgs_pairs=0, bulk=0.0, viscosity=50.0)
    parts = block_particles(s, (4, 4, 4), (12, 12, 12), vel=(0.3, -0.2, 0.1))
    sc = SceneArrays(parts)
    o = orc.OracleSim(s, sc.scene)
    dt = 1e-3
    o.debug_step_until(dt, "packing")
    before = o.read_grid_vel().copy()
    o.debug_step_until(dt, "viscosity")
    after = o.read_grid_vel()
    _, kind = o.read_pressure()
    core = np.nonzero(kind == 2)[0]
    assert core.size > 50
    # the grid field is uniform up to the P2G truncation (~1e-5 here); a blend
    # of weight 0.65 toward anything else would move it by ~0.1
    spread = np.abs(before[core, :3] - before[core, :3].mean(axis=0)).max()
    assert spread < 2e-4
    assert np.abs(after[core, :3] - before[core, :3]).max() <= 2 * spread + 4.0 / FP


def test_viscosity_blends_toward_neighbours(orc):
    """A shear layer: the blend moves each cell toward the mass-weighted
    neighbour average by alpha * weight_sum (shader.rs:1110-1162)."""
    s = box_settings(n=16, rbgs_pairs=0, bulk=0.0, viscosity=8.0)
    top = block_particles(s, (4, 8, 4), (12, 12, 12), vel=(1.0, 0.0, 0.0), seed=1)
    bot = block_particles(s, (4, 4, 4), (12, 8, 12), vel=(-1.0, 0.0, 0.0), seed=2)
    sc = SceneArrays(np.concatenate([top, bot]))
    o = orc.OracleSim(s, sc.scene)
    dt = 1e-3
    o.debug_step_until(dt, "packing")
    before = o.read_grid_vel().copy()
    o.debug_step_until(dt, "viscosity")
    after = o.read_grid_vel()
    # nodes 7 and 9 sit either side of node 8, which straddles the interface
    up = cell_index(s, 8, 9, 8)
    down = cell_index(s, 8, 7, 8)
    assert abs(before[cell_index(s, 8, 8, 8), 0]) < 0.2
    assert before[up, 0] > 0.3 and before[down, 0] < -0.3
    assert after[up, 0] < before[up, 0] - 0.01
    assert after[down, 0] > before[down, 0] + 0.01
    # the lateral neighbours of a cell deep in the top block move with it
    deep = cell_index(s, 8, 11, 8)
    assert after[deep, 0] > 0.0


def test_residual_leaves_out_cells_held_at_the_clamp(orc):
    """A block falling at 20 units/s into a cup: its impact drives cells to the
    pressure clamp (|p| = FP_VALUE_LIMIT - 1 = 8191).  The reported residual is
    ||b - A p|| / ||b|| of the system rebuilt here from the oracle's rows, over
    the fluid cells whose pressure is not at the clamp, and the solve reaches
    the 1e-4 tolerance; with the held cells included the residual could not."""
    s = box_settings(n=16, cup=True, tolerance=1e-4)
    s.rbgs_max_iters = 4000
    parts = block_particles(s, (4, 3, 4), (12, 9, 12), vel=(0.0, -20.0, 0.0))
    sc = SceneArrays(parts)
    o = orc.OracleSim(s, sc.scene)
    o.debug_step_until(1e-3, "rbgs")
    st = o.solve_stats()
    assert st.converged == 1 and st.iterations < 4000
    wts, count, c = o.read_system()
    p_fp, kind = o.read_pressure()
    p = p_fp.astype(np.float64) / FP
    fluid = np.nonzero((kind >= 1) & (kind <= 3) & (count > 0))[0]
    held = fluid[np.abs(p[fluid]) == 8191.0]
    assert held.size > 10, "the impact should drive cells to the clamp"
    n = s.grid_dims[0]
    off = [-1, 1, -n, n, -n * n, n * n]
    fluid_set = set(int(i) for i in fluid)

    def residual(rows):
        rr = bb = 0.0
        for i in rows:
            ap = float(count[i]) * p[i]
            for d in range(6):
                j = int(i) + off[d]
                if wts[i, d] > 0 and j in fluid_set:
                    ap -= float(wts[i, d]) * p[j]
            b = -float(c[i])
            rr += (b - ap) ** 2
            bb += b * b
        return (rr / bb) ** 0.5

    r_free = residual(fluid[np.abs(p[fluid]) != 8191.0])
    assert abs(r_free - st.final_residual) <= 1e-9 * r_free
    assert r_free <= 1e-4 < residual(fluid)
    o.close()
