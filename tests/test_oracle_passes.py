"""Closed-form checks of single oracle passes on hand-built grid states.

Each pass runs alone on a state written directly into the oracle (grid
velocity, grid lanes, particles), through `OracleSim.debug_run_pass`, and is
compared with what the mathematics fixes for that state -- not with a second
transcription of the pass:

- G2P (shader.rs:1683-1881): an affine grid field v = A x + b comes back as
  v_p = A x_p + b and C = A for fully supported particles (APIC exactness of
  the quadratic B-spline: zero first moment, second moment dx^2/4 per axis);
  the J update is J exp(dt tr A); a particle without grid support moves
  ballistically, scaled by the two preserve blends' closed forms; inside a
  porous bed the Darcy and inertial drag factors divide the velocity.
- boundary_project (shader.rs:1630-1679) against a planar SDF: the inward
  normal component is removed and the tangential part shrinks by Coulomb
  friction mu |v_n|.
- classify (shader.rs:1181-1367): the kind rules on a block next to a solid
  wall, the ghost mirror (a quiescent wall adds no divergence, a wall-normal
  flow adds v/dx), and the compression / expansion / overpack targets of
  shader.rs:597-619.
- packing (shader.rs:1566-1626): a linear overpack ramp gives the constant
  gradient K * slope, with the air (zero) and off-grid (mirror) rules.
- the post-projection divergence metric (no reference pass) on the same
  fields.
A plausible slip (a transposed C, a dropped B, the wrong trace, a sign of the
normal, a friction coefficient, a target weight) fails one of them."""
from __future__ import annotations

import numpy as np
import pytest

from scenes import DX, SceneArrays, bed_field, box_settings, cell_index

FP = 2.0 ** 18
M0 = 0.25  # nominal particle mass (mod.rs:575)


def _centers(s):
    n = int(s.grid_dims[0])
    origin = -0.5 * float(s.bounds_size[0])
    c = origin + (np.arange(n) + 0.5) * DX
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)  # cell order x fastest


def _particles(rng, s, count, lo_cell=3, hi_cell=13, J=1.1):
    n = int(s.grid_dims[0])
    origin = -0.5 * float(s.bounds_size[0])
    pos = origin + (lo_cell + rng.random((count, 3)) * (hi_cell - lo_cell)) * DX
    p = np.zeros((count, 8), np.float32)
    p[:, 0:3] = pos
    p[:, 3] = J
    p[:, 4:7] = rng.normal(0.0, 0.5, size=(count, 3))
    p[:, 7] = M0
    assert hi_cell + 2 <= n
    return p


def _sim(orc, s, parts, **scene_kw):
    sc = SceneArrays(parts, **{k: v for k, v in scene_kw.items() if k in ("bed_lookup", "bed_records")})
    if "sdf" in scene_kw:
        sc.sdf = np.ascontiguousarray(scene_kw["sdf"], np.float32)
        sc.cls = np.ascontiguousarray(scene_kw["cls"], np.uint8)
        import ctypes as C

        sc.scene.sdf = sc.sdf.ctypes.data_as(C.POINTER(C.c_float))
        sc.scene.sdf_class = sc.cls.ctypes.data_as(C.POINTER(C.c_uint8))
    return orc.OracleSim(s, sc.scene), sc


def _nodes(s):
    """Grid node positions as P2G/G2P place them: origin + i dx (shader.rs:975-988)."""
    return _centers(s) - 0.5 * DX


def _affine_field(s, A, b, mass):
    x = _nodes(s)
    gv = np.zeros((x.shape[0], 4), np.float32)
    gv[:, 0:3] = x @ A.T + b
    gv[:, 3] = mass
    return gv


A_TEST = np.array([[0.8, -1.3, 0.4], [2.1, -0.5, 0.9], [-0.7, 1.6, 1.1]])
B_TEST = np.array([0.3, -0.6, 0.45])


def test_g2p_reproduces_an_affine_field_exactly(orc):
    """shader.rs:1729-1755: v_p = A x_p + b and C = A (columns C0 = dv/dx ...)."""
    s = box_settings(n=16, rbgs_pairs=0)
    rng = np.random.default_rng(11)
    parts = _particles(rng, s, 40)
    o, _keep = _sim(orc, s, parts)
    o.debug_set_grid_vel(_affine_field(s, A_TEST, B_TEST, 10.0 * M0))
    dt = 1e-3
    o.debug_run_pass(dt, "g2p")
    got = o.read_particles().astype(np.float64)
    aff = o.read_affine().astype(np.float64)
    x0 = parts[:, 0:3].astype(np.float64)
    v_want = x0 @ A_TEST.T + B_TEST
    assert np.abs(got[:, 4:7] - v_want).max() < 2e-5 * (1 + np.abs(v_want).max())
    for col in range(3):
        c_got = aff[:, 4 * col:4 * col + 3]
        assert np.abs(c_got - A_TEST[:, col]).max() < 2e-4, col
    assert np.all(aff[:, 3] == 0.0)  # phase kept
    # x += v dt (no contact: the box has no obstacles)
    assert np.abs(got[:, 0:3] - (x0 + got[:, 4:7] * dt)).max() < 1e-6
    # J <- J exp(dt tr C) with full support (shader.rs:1851-1855)
    J_want = 1.1 * np.exp(dt * np.trace(A_TEST))
    assert np.abs(got[:, 3] - J_want).max() < 2e-6


def test_g2p_j_update_clamps(orc):
    """exp argument clamped to +-0.35, J to [0.4, 2] before and after (shader.rs:1851-1855)."""
    s = box_settings(n=16, rbgs_pairs=0)
    rng = np.random.default_rng(12)
    parts = _particles(rng, s, 8, J=1.5)
    parts[4:, 3] = 3.0  # clamped to 2 before the update
    o, _keep = _sim(orc, s, parts)
    A = np.diag([1.0, 1.0, 1.0])  # tr = 3
    o.debug_set_grid_vel(_affine_field(s, A, np.zeros(3), 10.0 * M0))
    dt = 0.2  # dt tr C = 0.6 > 0.35
    o.debug_run_pass(dt, "g2p")
    J = o.read_particles()[:, 3].astype(np.float64)
    assert np.allclose(J[:4], min(1.5 * np.exp(0.35), 2.0), rtol=2e-6)
    assert np.allclose(J[4:], 2.0, rtol=0)
    A = -A
    o2, _k2 = _sim(orc, s, parts)
    o2.debug_set_grid_vel(_affine_field(s, A, np.zeros(3), 10.0 * M0))
    o2.debug_run_pass(dt, "g2p")
    J = o2.read_particles()[:, 3].astype(np.float64)
    assert np.allclose(J[:4], 1.5 * np.exp(-0.35), rtol=2e-6)
    assert np.allclose(J[4:], 2.0 * np.exp(-0.35), rtol=2e-6)


def test_g2p_unsupported_particle_is_ballistic(orc):
    """No grid mass: the weak-support blend (gain 1.35, cap 0.995) then the
    airborne blend (gain 0.98, cap 0.98) leave v = (0.995 * 0.02 + 0.98) *
    (v + g dt) and C scaled by (1 - 0.995 * 0.75)(1 - 0.98 * 0.65); J is kept
    (shader.rs:1793-1827, 1854)."""
    s = box_settings(n=16, rbgs_pairs=0)
    rng = np.random.default_rng(13)
    parts = _particles(rng, s, 12)
    aff = np.zeros((len(parts), 12), np.float32)
    aff[:, 0:3] = rng.normal(size=(len(parts), 3))
    sc = SceneArrays(parts, aff)
    o = orc.OracleSim(s, sc.scene)
    o.debug_set_grid_vel(np.zeros((s.num_cells, 4), np.float32))
    dt = 1e-3
    o.debug_run_pass(dt, "g2p")
    got = o.read_particles().astype(np.float64)
    ball = parts[:, 4:7].astype(np.float64)
    ball[:, 1] += s.gravity * dt
    f = 0.995 * (1 - 0.98) + 0.98
    assert np.abs(got[:, 4:7] - f * ball).max() < 1e-5
    assert np.allclose(got[:, 3], 1.1, rtol=1e-7)
    # no support: the gathered C is 0 before the scaling
    assert np.all(o.read_affine()[:, 0:3] == 0.0)


def test_g2p_darcy_and_inertial_drag_inside_the_bed(orc):
    """Bed overlap 1: v = v_rel / ((1 + nu dt / k)(1 + 0.35 |v_rel| dt / dx)) and
    C scaled by the same factor (shader.rs:1829-1849), bed at rest."""
    s = box_settings(n=16, rbgs_pairs=0)
    s.has_bed_field = 1
    k = 2.0e-9
    lookup, recs = bed_field(s, (0, 0, 0), (16, 16, 16), porosity=0.4, permeability=k)
    rng = np.random.default_rng(14)
    parts = _particles(rng, s, 30)
    o, _keep = _sim(orc, s, parts, bed_lookup=lookup, bed_records=recs)
    o.debug_set_grid_vel(_affine_field(s, A_TEST, B_TEST, 10.0 * M0))
    dt = 1e-3
    o.debug_run_pass(dt, "g2p")
    got = o.read_particles().astype(np.float64)
    aff = o.read_affine().astype(np.float64)
    v_apic = parts[:, 0:3].astype(np.float64) @ A_TEST.T + B_TEST
    dd = 1.0 / (1.0 + float(s.water_viscosity_m2_s) / k * dt)
    idd = 1.0 / (1.0 + np.linalg.norm(v_apic, axis=1) * dt / DX * 0.35)
    fac = dd * idd
    assert np.abs(got[:, 4:7] - v_apic * fac[:, None]).max() < 3e-5
    for col in range(3):
        assert np.abs(aff[:, 4 * col:4 * col + 3] - A_TEST[:, col] * fac[:, None]).max() < 2e-4


def _plane_sdf(s, y0):
    r = int(s.sdf_res)
    bs = float(s.bounds_size[1])
    yc = (np.arange(r) + 0.5) / r * bs - 0.5 * bs
    sdf = np.broadcast_to((yc - y0)[None, :, None], (r, r, r)).astype(np.float32)
    return np.ascontiguousarray(sdf)  # [z][y][x]


def test_boundary_project_against_a_plane(orc):
    """Cells with sdf(center) < 0.05 and v.y < 0 lose v.y and their tangential
    speed shrinks by min(0.3 |v.y|, |v_t|) (shader.rs:1630-1679, restitution 0)."""
    s = box_settings(n=16, rbgs_pairs=0)
    origin = -0.5 * float(s.bounds_size[1])
    y0 = origin + 7.3 * DX
    o, _keep = _sim(orc, s, np.zeros((0, 8), np.float32), sdf=_plane_sdf(s, y0),
                    cls=np.zeros(s.num_cells, np.uint8))
    rng = np.random.default_rng(21)
    gv = np.zeros((s.num_cells, 4), np.float32)
    gv[:, 0:3] = rng.normal(0.0, 2.0, size=(s.num_cells, 3))
    gv[:, 3] = 1.0
    o.debug_set_grid_vel(gv)
    o.debug_run_pass(1e-3, "boundary")
    got = o.read_grid_vel().astype(np.float64)
    x = _centers(s)
    inner = np.all((x > origin + 2.5 * DX) & (x < -origin - 2.5 * DX), axis=1)  # outside the box margin
    v = gv[:, 0:3].astype(np.float64)
    sd = x[:, 1] - y0
    hit = (sd < 0.05) & (v[:, 1] < 0.0)
    assert hit[inner].sum() > 100 and (~hit[inner]).sum() > 100
    want = v.copy()
    vt = v[hit].copy()
    vt[:, 1] = 0.0
    L = np.linalg.norm(vt, axis=1)
    fi = np.minimum(0.3 * np.abs(v[hit, 1]), L)
    want[hit] = vt * (1.0 - fi / L)[:, None]
    assert np.abs(got[inner, 0:3] - want[inner]).max() < 1e-5
    assert np.all(got[:, 3] == 1.0)


def _classify_scene(orc, n=8, vel=(0.0, 0.0, 0.0)):
    """A fluid block ix 1..5, iy/iz 1..5 beside a solid wall at ix = 0 (the
    solid cells carry mass, so the wall neighbours stay interior)."""
    s = box_settings(n=n, rbgs_pairs=0)
    cls = np.zeros(s.num_cells, np.uint8)
    mass = np.zeros(s.num_cells)
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                c = cell_index(s, ix, iy, iz)
                if ix == 0:
                    cls[c] = 1
                    mass[c] = 1.0
                elif ix <= 5 and 1 <= iy <= 5 and 1 <= iz <= 5:
                    mass[c] = 1.0
    o, keep = _sim(orc, s, np.zeros((0, 8), np.float32), sdf=np.full(int(s.sdf_res) ** 3, 999.0, np.float32),
                   cls=cls)
    gv = np.zeros((s.num_cells, 4), np.float32)
    gv[:, 0:3] = vel
    gv[:, 3] = mass
    o.debug_set_grid_vel(gv)
    return s, o, keep, cls, mass


def _kinds_expected(s, cls, mass):
    n = int(s.grid_dims[0])
    occ = 0.1 * M0
    k = np.zeros(s.num_cells, np.int64)
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                c = cell_index(s, ix, iy, iz)
                if cls[c]:
                    k[c] = 4
                elif mass[c] <= occ:
                    k[c] = 0
                else:
                    air = False
                    for d in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                        x, y, z = ix + d[0], iy + d[1], iz + d[2]
                        if not (0 <= x < n and 0 <= y < n and 0 <= z < n) or mass[cell_index(s, x, y, z)] <= occ:
                            air = True
                    k[c] = 1 if air else 2
    return k


def test_classify_kinds_and_quiescent_wall(orc):
    s, o, _keep, cls, mass = _classify_scene(orc)
    lanes = np.zeros((6, s.num_cells), np.int32)
    cv = DX ** 3
    lanes[4] = lanes[5] = int(cv * FP)  # rest = current: no volume target
    o.debug_set_grid_lanes(lanes)
    o.debug_run_pass(1e-3, "classify")
    got = o.debug_read_grid_lanes()
    assert np.array_equal(got[3], _kinds_expected(s, cls, mass))
    assert np.all(got[0] == 0)  # pressure zeroed (shader.rs:1187)
    assert np.all(got[1] == 0)  # v = 0: the ghost mirror adds no divergence at the wall
    m = o.read_metrics()
    assert m[1] == np.count_nonzero((got[3] >= 1) & (got[3] <= 3))


def test_classify_ghost_mirror_and_targets(orc):
    c = 0.7
    s, o, _keep, cls, mass = _classify_scene(orc, vel=(c, 0.0, 0.0))
    n = int(s.grid_dims[0])
    cv = np.float32(DX) ** 3
    lanes = np.zeros((6, s.num_cells), np.int32)
    # per-cell volume states by iz: compression, expansion, overpack, none
    states = {1: (1.0, 0.9), 2: (1.0, 1.1), 3: (1.5, 1.5), 4: (1.0, 1.0), 5: (0.0, 0.0)}
    target_of = {}
    for iz in range(n):
        rf, cf = states.get(iz, (1.0, 1.0))
        rl, cl = int(rf * cv * FP), int(cf * cv * FP)  # the lanes hold fixed point
        for iy in range(n):
            for ix in range(n):
                ci = cell_index(s, ix, iy, iz)
                lanes[4, ci] = rl
                lanes[5, ci] = cl
        rest, cur = rl / FP, cl / FP
        if rest <= max(float(cv) * 1e-5, DX ** 3 * 0.125 * 1e-3):
            target_of[iz] = 0.0
            continue
        j = cur / rest
        t = np.clip(1 - j, 0, 0.75) * 32.0 - np.clip(j - 1, 0, 0.75) * 2.0
        target_of[iz] = max(t, max(max(rest, cur) / float(cv) - 1.2, 0.0) * 32.0)
    # the expansion term -2 (j - 1) never survives the overpack max, which is
    # at least 0 (shader.rs:604-616): an expanded cell's target is 0
    assert target_of[1] > 3.0 and target_of[2] == 0.0 and target_of[3] > 9.0 and target_of[5] == 0.0
    o.debug_set_grid_lanes(lanes)
    o.debug_run_pass(1e-3, "classify")
    got = o.debug_read_grid_lanes()
    kind = got[3]
    checked = 0
    for iz in range(1, 6):
        for iy in range(1, 6):
            for ix in range(1, 6):
                ci = cell_index(s, ix, iy, iz)
                assert kind[ci] in (1, 2)
                # x faces: solid wall at ix = 0 mirrors -c; air at ix = 6 still reads c
                div = c / DX if ix == 1 else 0.0
                want = (div - target_of[iz]) * FP
                assert abs(got[1, ci] - want) <= 2e-5 * abs(want) + 8, (ix, iy, iz, got[1, ci], want)
                checked += 1
    assert checked == 125
    # the metric slot 0 holds max|div| (no target): c / dx, encoded x1024
    assert abs(int(o.read_metrics()[0]) - c / DX * 1024) <= 1


def test_post_projection_divergence_metric(orc):
    """Slots 5/6: max|div| and max|div - target| of the current grid velocity
    over classify's fluid cells (here the classify field itself)."""
    c = 0.7
    s, o, _keep, cls, mass = _classify_scene(orc, vel=(c, 0.0, 0.0))
    cv = np.float32(DX) ** 3
    lanes = np.zeros((6, s.num_cells), np.int32)
    lanes[4] = int(cv * FP)
    lanes[5] = int(np.float32(0.9) * cv * FP)  # compression: target 32 (1 - j), j ~ 0.9
    tgt = 32.0 * (1.0 - float(lanes[5, 0]) / float(lanes[4, 0]))
    o.debug_set_grid_lanes(lanes)
    o.debug_run_pass(1e-3, "classify")
    o.debug_clear_metrics()
    o.debug_run_pass(1e-3, "post_divergence")
    m = o.read_metrics()
    assert abs(int(m[5]) - c / DX * 1024) <= 1
    # the wall cells: |c/dx - tgt| ~ 0.8; the others |0 - tgt| ~ 3.2
    assert abs(int(m[6]) - max(abs(c / DX - tgt), tgt) * 1024) <= 1
    # a divergence-free field after "projection": only the targets remain
    gv = np.zeros((s.num_cells, 4), np.float32)
    gv[:, 3] = mass
    o.debug_set_grid_vel(gv)
    o.debug_clear_metrics()
    o.debug_run_pass(1e-3, "post_divergence")
    m = o.read_metrics()
    assert m[5] == 0 and abs(int(m[6]) - tgt * 1024) <= 1


def test_packing_ramp_gives_a_constant_gradient(orc):
    """p_pack = K max(packed - 1, 0) (shader.rs:1566-1583), v -= dt grad p with
    central differences; an air neighbour reads 0 (shader.rs:1595-1611)."""
    n = 8
    s = box_settings(n=n, rbgs_pairs=0, bulk=900.0)
    mass = np.zeros(s.num_cells)
    for iz in range(n):
        for iy in range(n):
            for ix in range(7):  # ix = 7 stays air
                mass[cell_index(s, ix, iy, iz)] = 1.0
    o, _keep = _sim(orc, s, np.zeros((0, 8), np.float32), sdf=np.full(int(s.sdf_res) ** 3, 999.0, np.float32),
                    cls=np.zeros(s.num_cells, np.uint8))
    gv = np.zeros((s.num_cells, 4), np.float32)
    gv[:, 3] = mass
    o.debug_set_grid_vel(gv)
    cv = np.float32(DX) ** 3
    lanes = np.zeros((6, s.num_cells), np.int32)
    frac = np.zeros(s.num_cells)
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                ci = cell_index(s, ix, iy, iz)
                frac[ci] = 1.0 + 0.05 * ix
                lanes[4, ci] = lanes[5, ci] = int(frac[ci] * cv * FP)
    o.debug_set_grid_lanes(lanes)
    dt = 1e-3
    o.debug_run_pass(dt, "classify")
    o.debug_run_pass(dt, "packing")
    got = o.read_grid_vel().astype(np.float64)
    K = 900.0

    def p(ix):  # K (packed - 1) from the fixed-point lane, stored as fixed point
        packed = int((1.0 + 0.05 * ix) * cv * FP) / FP / float(cv)
        return np.trunc(K * max(packed - 1.0, 0.0) * FP) / FP
    for iz in range(1, n - 1):
        for iy in range(1, n - 1):
            for ix in range(7):
                ci = cell_index(s, ix, iy, iz)
                pm = p(ix) if ix == 0 else p(ix - 1)  # off grid: mirror
                pp = 0.0 if ix == 6 else p(ix + 1)  # air neighbour: 0
                want = -dt * 0.5 / DX * (pp - pm)
                assert abs(got[ci, 0] - want) <= 1e-5 * abs(want) + 1e-6, (ix, got[ci, 0], want)
                assert abs(got[ci, 1]) < 1e-6 and abs(got[ci, 2]) < 1e-6
