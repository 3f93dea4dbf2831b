"""The bed grains' own motion (bed_dynamics, shader.rs:1991-2238) and the
water's reaction on them (bed_impulse_*, shader.rs:458-522), checked in the
oracle on hand-built states against special cases whose outcome the rules fix:

- a lone anchored grain in dry air only damps: v <- v (1 - 8.5 dt), x += v dt;
- a reaction impulse is divided by max(m, m0/4) and capped in magnitude at
  min(0.012 v_cap, 0.02 dx/dt), and its lanes are consumed;
- a lone suspended grain blends toward the settling velocity (0, -1.10, 0)
  with weight 8 dt, and redeposits (phase 1) once stamped grains surround it;
- the packing contact pushes a grain out of a neighbour closer than 1.78 grain
  radii by min(0.65 penetration, 0.22 dx) and removes the approach velocity;
- a linear pressure ramp drives an anchored, wetted grain down the gradient,
  and a neighbour whose kind lane reads solid halves the gradient (mirror);
- the Darcy damping in project_pressure hands the grains alpha = 0.04 of the
  momentum it takes from the water, split over the stamped cells.

A slip (a missing damping factor, the impulse not divided by the mass, the
wrong clamp, a packing push toward the neighbour, the mirror taken on the
wrong side, the reaction's sign or alpha) fails one of them."""
from __future__ import annotations

import numpy as np
import pytest

import coffee_sim_amd as cs
from scenes import DX, box_settings

FP = 2.0 ** 18
M0 = 0.25
UPM = float(np.float32(np.float32(2.75) - np.float32(-3.02)) / np.float32(0.10))  # sim units per metre
VCAP = 1.5 * UPM
DT = 1e-3
N = 16


def _cell_center(ix, iy, iz):
    o = -0.5 * N * DX
    return np.array([o + (ix + 0.5) * DX, o + (iy + 0.5) * DX, o + (iz + 0.5) * DX], np.float64)


def grain_scene(grains, sat=0.0, compaction=0.0):
    """A 16^3 box with a few water particles in a corner and the given grains:
    dicts with pos, vel, phase, rest (default pos), mass (default M0)."""
    s = box_settings(n=N, rbgs_pairs=0)
    nb = len(grains)
    part = np.zeros((nb, 8), np.float32)
    aff = np.zeros((nb, 12), np.float32)
    ext = np.zeros((nb, 8), np.float32)
    for i, g in enumerate(grains):
        part[i, 0:3] = g["pos"]
        part[i, 3] = 1.0
        part[i, 4:7] = g.get("vel", (0.0, 0.0, 0.0))
        part[i, 7] = g.get("mass", M0)
        aff[i, 3] = g.get("phase", 1.0)
        aff[i, 4:7] = g.get("rest", g["pos"])
        ext[i] = [0.0, 0.4, 1e-9, compaction, 0.15, 0.0, 0.0, sat]
    water = np.zeros((8, 8), np.float32)
    water[:, 0:3] = _cell_center(1, 1, 1) + 0.01 * np.arange(8)[:, None]
    water[:, 3] = 1.0
    water[:, 7] = M0
    sc = cs.Scene()
    wa = np.zeros((8, 12), np.float32)
    sc.particles = cs._fptr(water)
    sc.affine = cs._fptr(wa)
    sc.num_particles = 8
    sc._keep = [water, wa]
    cs.attach_bed(sc, {"particles": part, "affine": aff, "extracts": ext})
    return s, sc


def run_dynamics(orc, s, sc, delta=None, vel4=None, lanes=None):
    o = orc.OracleSim(s, sc)
    o.debug_run_pass(DT, "bed_stamp")
    if vel4 is not None:
        o.debug_set_grid_vel(vel4)
    if lanes is not None:
        o.debug_set_grid_lanes(lanes)
    if delta is not None:
        o.debug_set_bed_delta(delta)
    o.debug_run_pass(DT, "bed_dynamics")
    p, a = o.read_bed_particles()
    return o, p.astype(np.float64), a.astype(np.float64)


def test_lone_anchored_grain_in_dry_air_only_damps(orc):
    x0 = _cell_center(8, 8, 8)
    v0 = np.array([0.5, -0.5, 0.3])
    s, sc = grain_scene([dict(pos=x0, vel=v0)])
    o, p, a = run_dynamics(orc, s, sc)
    v = v0 * (1.0 - 8.5 * DT)
    assert np.abs(p[0, 4:7] - v).max() < 1e-6
    assert np.abs(p[0, 0:3] - (x0 + v * DT)).max() < 1e-6
    assert np.array_equal(a[0, 4:7], x0.astype(np.float32))  # rest unmoved: no lateral excess, no compaction
    assert a[0, 3] == 1.0 and p[0, 7] == np.float32(M0)
    # the compaction marker takes the sag below rest in cells
    ext = o.read_bed_extracts()
    assert abs(ext[0, 3] - 0.5 * (1.0 - 8.5 * DT) * DT / DX) < 1e-6


@pytest.mark.parametrize("mass, imp, capped", [(0.01, (0.01, -0.004, 0.002), False), (M0, (1.0, 0.5, 0.0), True)])
def test_reaction_impulse_divided_by_mass_and_capped(orc, mass, imp, capped):
    x0 = _cell_center(8, 8, 8)
    s, sc = grain_scene([dict(pos=x0, mass=mass)])
    delta = np.zeros((4, 1), np.int32)
    delta[0, 0] = 12345  # the water credit lane is not bed_dynamics' to consume
    fx = np.trunc(np.array(imp) * FP).astype(np.int32)
    delta[1:4, 0] = fx
    o, p, _ = run_dynamics(orc, s, sc, delta=delta)
    brv = fx / FP / max(mass, M0 * 0.25)
    max_rs = min(VCAP * 0.012, DX / DT * 0.02)
    r = np.linalg.norm(brv)
    assert (r > max_rs) == capped
    if capped:
        brv = brv * (max_rs / r)
    assert np.abs(p[0, 4:7] - brv * (1.0 - 8.5 * DT)).max() < 2e-6 * max(1.0, np.abs(brv).max())
    d = o.debug_read_bed_delta()
    assert d[0, 0] == 12345 and (d[1:, 0] == 0).all()


def test_lone_suspended_grain_settles(orc):
    x0 = _cell_center(8, 9, 8)
    v0 = np.array([0.4, 0.2, -0.1])
    s, sc = grain_scene([dict(pos=x0, vel=v0, phase=2.0)], compaction=0.4)
    o, p, a = run_dynamics(orc, s, sc)
    drag = 8.0 * DT
    unsup = v0 + np.array([0.0, s.gravity * DT * 0.85, 0.0])
    sv = unsup * (1.0 - drag) + np.array([0.0, -1.10, 0.0]) * drag
    assert np.abs(p[0, 4:7] - sv).max() < 1e-5
    assert np.abs(p[0, 0:3] - (x0 + sv * DT)).max() < 1e-6
    assert a[0, 3] == 2.0  # nothing around it: stays suspended
    assert np.array_equal(a[0, 4:7], p[0, 0:3])  # rest follows a suspended grain
    assert abs(o.read_bed_extracts()[0, 3] - 0.4 * 0.985) < 1e-7


def test_suspended_grain_redeposits_among_stamped_grains(orc):
    """Grain 0 suspended just above an anchored grain 1 (the higher index owns
    the shared stamp): its B-spline stencil sees grain 1's cells, so it is
    anchored by its neighbours and redeposits while settling."""
    x0 = _cell_center(8, 9, 8)
    x1 = _cell_center(8, 8, 8)
    s, sc = grain_scene([dict(pos=x0, phase=2.0), dict(pos=x1)])
    _, _, a = run_dynamics(orc, s, sc)
    assert a[0, 3] == 1.0 and a[1, 3] == 1.0


def test_packing_pushes_out_of_a_close_neighbour(orc):
    r = 0.62 * DX
    min_sep = 1.78 * r
    x1 = _cell_center(9, 8, 8)
    for d, vx in ((1.0 * DX, 0.2), (0.5 * DX, 0.0), (1.2 * DX, 0.0)):
        x0 = x1 - np.array([d, 0.0, 0.0])
        s, sc = grain_scene([dict(pos=x0, vel=(vx, 0.0, 0.0)), dict(pos=x1)])
        _, p, _ = run_dynamics(orc, s, sc)
        v = vx * (1.0 - 8.5 * DT)
        moved = x0[0] + v * DT
        closest = x1[0] - moved
        push = min(0.65 * (min_sep - closest), 0.22 * DX) if closest < min_sep else 0.0
        assert abs(p[0, 0] - (moved - push)) < 2e-6, d  # pushed away from the neighbour
        assert p[0, 1] == np.float32(x0[1]) and p[0, 2] == np.float32(x0[2])
        assert p[0, 4] == (0.0 if push > 0.0 else np.float32(v))  # approach velocity removed
        # the neighbour (it owns every stamped cell around itself) sees no one
        assert np.array_equal(p[1, 0:3], x1.astype(np.float32))


def _wet_grid(mass):
    vel4 = np.zeros((N ** 3, 4), np.float32)
    vel4[:, 3] = mass
    return vel4


@pytest.mark.parametrize("solid_plus_x", [False, True])
def test_pressure_ramp_drives_a_wetted_grain_down_the_gradient(orc, solid_plus_x):
    ic = (8, 8, 8)
    x0 = _cell_center(*ic)
    s, sc = grain_scene([dict(pos=x0)])
    mass = 0.3
    slope = 0.02  # pressure per cell along +x
    lanes = np.zeros((6, N, N, N), np.int32)  # [lane][z][y][x]
    lanes[0] = np.trunc(slope * np.arange(N)[None, None, :] * FP).astype(np.int32)
    if solid_plus_x:
        lanes[3, ic[2], ic[1], ic[0] + 1] = 4  # CELL_SOLID read through the kind lane: mirror
    _, p, _ = run_dynamics(orc, s, sc, vel4=_wet_grid(mass), lanes=lanes.reshape(6, -1))
    ph = lanes[0, ic[2], ic[1], ic[0]] / FP
    pxp = ph if solid_plus_x else lanes[0, ic[2], ic[1], ic[0] + 1] / FP
    pxm = lanes[0, ic[2], ic[1], ic[0] - 1] / FP
    gx = 0.5 / DX * (pxp - pxm)
    wl = min(mass / (M0 * 1.5), 1.0)
    mob = min(0.30 + wl * 0.62, 1.0)
    sf = min(max((x0[1] + 3.0) / 3.5, 0.12), 1.0)
    pl = min(mass * 0.25, 1.0)
    damp = 1.0 - 8.5 * DT
    ed = min(2.0 * wl * DT, 0.08)
    vx = -np.sign(gx) * min(abs(gx), VCAP / DT) * pl * 0.08 * 8.0 * mob * sf * DT * damp
    vy = -(0.95 * (1 - wl) + 0.35 * wl) * ed * damp
    assert abs(p[0, 4] - vx) < 1e-6 and abs(p[0, 5] - vy) < 1e-6 and abs(p[0, 6]) < 1e-9
    assert vx < 0.0  # down the gradient


def test_project_reaction_returns_alpha_of_the_darcy_momentum(orc):
    """Grains at rest on every cell of an 4^3 block; the bed cells' water moves
    at v0.  project_pressure with zero pressure damps it toward the (resting)
    matrix, and the grains' impulse lanes gain alpha times the momentum the
    water lost, up to one fixed-point unit per share."""
    cells = [(x, y, z) for z in range(6, 10) for y in range(6, 10) for x in range(6, 10)]
    grains = [dict(pos=_cell_center(*c)) for c in cells]
    s, sc = grain_scene(grains)
    o = orc.OracleSim(s, sc)
    o.debug_run_pass(DT, "bed_stamp")
    look = o.read_bed_lookup()
    bedcells = np.nonzero(look >= 0)[0]
    v0 = np.array([0.3, -0.5, 0.2], np.float32)
    mass = 0.6
    vel4 = np.zeros((N ** 3, 4), np.float32)
    vel4[bedcells, 0:3] = v0
    vel4[bedcells, 3] = mass
    lanes = np.zeros((6, N ** 3), np.int32)
    lanes[3, bedcells] = 3  # CELL_BED
    o.debug_set_grid_vel(vel4)
    o.debug_set_grid_lanes(lanes)
    o.debug_run_pass(DT, "project")
    after = o.read_grid_vel()[bedcells, 0:3].astype(np.float64)
    lost = ((v0.astype(np.float64) - after) * mass).sum(axis=0)
    assert (np.abs(after) < np.abs(v0)).all()  # damped toward the resting matrix
    got = o.debug_read_bed_delta()[1:4].astype(np.float64).sum(axis=1) / FP
    shares = 27 * len(bedcells)
    assert np.abs(got - 0.04 * lost).max() <= shares / FP + 1e-6 * np.abs(lost).max(), (got, 0.04 * lost)
    assert (np.sign(got) == np.sign(v0)).all()  # pushed along the flow
