"""Readback statistics of the reference's behavioural scenarios, restated
over the particle arrays (physics_tests.rs:53-1092).

Particles whose mass is at most 0.1 nominal mass are inactive; the particle
volume is dx^3 / 4 per nominal mass.  Sums are taken in float64; the
reference's thresholds leave far more room than the difference to its
float32 sums."""
from __future__ import annotations

import numpy as np

import coffee_sim_amd as cs

FRAME = 1.0 / 60.0
NOMINAL_MASS = 80.0 / 320.0
INACTIVE = NOMINAL_MASS * 0.1


def frames(sim, n: int) -> None:
    for _ in range(n):
        sim.step_frame(FRAME)


def _water(sim):
    """The water particles (the bed's live in their own array)."""
    return sim.read_particles().astype(np.float64)


def _all_particles(sim):
    """Bed particles first, then water, as the reference's particle buffer orders them."""
    nb = sim.counts()[1]
    w = _water(sim)
    if nb == 0:
        return w
    return np.concatenate([sim.read_bed_particles()[0].astype(np.float64), w])


def _dx(sim) -> float:
    return float(sim.settings.bounds_size[0]) / int(sim.settings.grid_dims[0])


# ---- snapshots (physics_tests.rs:61-1083) -------------------------------------------------
def mass_snapshot(sim) -> dict:
    """readback_mass_snapshot (physics_tests.rs:53-139): particle mass over the
    whole buffer and over the water, the bed-held water (BedExtract x), the
    water slots in use."""
    nw, nb, _ = sim.counts()
    w = _water(sim)
    bed = sim.read_bed_particles()[0].astype(np.float64) if nb else np.zeros((0, 8))
    held = float(sim.read_bed_extracts()[:, 0].astype(np.float64).sum()) if nb else 0.0
    water_mass = float(w[:, 7].sum())
    return {"active_particle_mass": water_mass + float(bed[:, 7].sum()), "active_water_particle_mass": water_mass,
            "bed_held_mass": held, "water_slots": nw}


def volume_snapshot(sim) -> dict:
    w = _water(sim)
    finite = bool(np.isfinite(w[:, [0, 1, 2, 3, 7]]).all())
    act = w[w[:, 7] > INACTIVE]
    pv = _dx(sim) ** 3 * 0.25
    scale = act[:, 7] / NOMINAL_MASS
    j = act[:, 3]
    n = max(len(act), 1)
    return {"all_finite": finite, "active_count": len(act), "active_mass": float(act[:, 7].sum()),
            "rest_volume": float((scale * pv).sum()), "current_volume": float((scale * pv * j).sum()),
            "mean_j": float(j.sum() / n), "min_j": float(j.min()) if len(act) else 0.0,
            "max_j": float(j.max()) if len(act) else 0.0}


def packing_snapshot(sim) -> dict:
    w = _water(sim)
    finite = bool(np.isfinite(w[:, [0, 1, 2, 3, 7]]).all())
    act = w[w[:, 7] > INACTIVE]
    gx, gy, gz = (int(v) for v in sim.settings.grid_dims)
    dx = _dx(sim)
    cell_v = dx ** 3
    origin = -0.5 * np.array([float(v) for v in sim.settings.bounds_size])
    g = (act[:, :3] - origin) / dx
    base = np.floor(g - 0.5)
    fx = g - base
    wts = [np.stack([0.5 * (1.5 - fx[:, a]) ** 2, 0.75 - (fx[:, a] - 1.0) ** 2, 0.5 * (fx[:, a] - 0.5) ** 2])
           for a in range(3)]
    rest_p = cell_v * 0.25 * act[:, 7] / NOMINAL_MASS
    cur_p = rest_p * np.clip(act[:, 3], 0.40, 2.00)
    rest = np.zeros(gx * gy * gz)
    cur = np.zeros(gx * gy * gz)
    b = base.astype(np.int64)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                cx, cy, cz = b[:, 0] + i, b[:, 1] + j, b[:, 2] + k
                ok = (cx >= 0) & (cy >= 0) & (cz >= 0) & (cx < gx) & (cy < gy) & (cz < gz)
                wgt = wts[0][i] * wts[1][j] * wts[2][k]
                cell = (cz * gy + cy) * gx + cx
                np.add.at(rest, cell[ok], (wgt * rest_p)[ok])
                np.add.at(cur, cell[ok], (wgt * cur_p)[ok])
    rf, cf = rest / cell_v, cur / cell_v
    packed = np.maximum(rf, cf)
    frac = (packed > 0.0) & (packed < 1.0)
    over = packed > 1.20
    return {"all_finite": finite, "active_count": len(act), "deposited_rest_volume": float(rest.sum()),
            "deposited_current_volume": float(cur.sum()), "max_rest_fraction": float(rf.max()),
            "max_current_fraction": float(cf.max()), "max_packed_fraction": float(packed.max()),
            "fractional_cell_count": int(frac.sum()),
            "max_fractional_fraction": float(packed[frac].max()) if frac.any() else 0.0,
            "overpacked_cell_count": int(over.sum()),
            "overpacked_volume_fraction": float((packed[over] - 1.20).sum())}


def velocity_snapshot(sim, y_min: float = -np.inf, y_max: float = np.inf) -> dict:
    w = _water(sim)
    finite = bool(np.isfinite(w[:, [1, 4, 5, 6, 7]]).all())
    act = w[(w[:, 7] > INACTIVE) & (w[:, 1] >= y_min) & (w[:, 1] <= y_max)]
    m, v = act[:, 7], act[:, 4:7]
    sp2 = (v ** 2).sum(1)
    mass = float(m.sum())
    n = max(len(act), 1)
    return {"all_finite": finite, "active_count": len(act), "active_mass": mass,
            "kinetic_energy": float((0.5 * m * sp2).sum()),
            "rms_speed": float(np.sqrt((m * sp2).sum() / max(mass, 1e-6))),
            "mean_speed": float(np.sqrt(sp2).sum() / n),
            "lateral_rms_speed": float(np.sqrt((m * (v[:, 0] ** 2 + v[:, 2] ** 2)).sum() / max(mass, 1e-6))),
            "max_speed": float(np.sqrt(sp2).max()) if len(act) else 0.0,
            "momentum": (m[:, None] * v).sum(0)}


def diag_snapshot(sim, bed_only: bool = False) -> dict:
    """readback_diag_snapshot (physics_tests.rs:937-1075) over the whole
    particle buffer, or over the bed particles (readback_bed_diag_snapshot)."""
    p = sim.read_bed_particles()[0].astype(np.float64) if bed_only else _all_particles(sim)
    finite = bool(np.isfinite(p[:, [0, 1, 2, 3, 7]]).all())
    act = p[p[:, 7] > INACTIVE]
    y, j = act[:, 1], act[:, 3]
    n = max(len(act), 1)
    ext = {f"{ax}_extent": float(act[:, i].max() - act[:, i].min()) if len(act) else 0.0
           for i, ax in enumerate("xyz")}
    return {"all_finite": finite, "total_mass": float(act[:, 7].sum()), "active_count": len(act),
            **ext, "y_mean": float(y.sum() / n),
            "mean_j": float(j.sum() / n), "min_j": float(j.min()) if len(act) else 0.0,
            "max_j": float(j.max()) if len(act) else 0.0}


# ---- the bed (physics_tests.rs:573-936, 1077-1090) -----------------------------------------
def bed_bounds_y() -> tuple[float, float]:
    """benchmark_bed_bounds_y (physics_tests.rs:1085-1090): the seated bed's
    bottom and top in world y."""
    b = cs.bed_config_seated_in_filter()
    return float(b.center[1] + b.bot_y), float(b.center[1] + b.top_y)


def filter_inner_radius(f, y_local: float) -> float:
    """FilterConfig::inner_radius_at_y (filter.rs:28-37)."""
    h = max(f.top_y - f.bot_y, 1e-6)
    t = min(max((y_local - f.bot_y) / h, 0.0), 1.0)
    return max(f.bot_radius + (f.top_radius - f.bot_radius) * t - f.thickness, f.hole_radius)


def bed_containment_snapshot(sim) -> dict:
    """readback_bed_filter_containment_snapshot (physics_tests.rs:608-693)."""
    f = cs.filter_config_default()
    p = sim.read_bed_particles()[0].astype(np.float64)
    finite = bool(np.isfinite(p[:, [0, 1, 2, 7]]).all())
    act = p[p[:, 7] > INACTIVE]
    r_bed = _dx(sim) * 0.62
    bot = f.center[1] + f.bot_y
    excess, clear = -np.inf, np.inf
    for x, y, z in act[:, :3]:
        y_local = min(max(y - f.center[1], f.bot_y), f.top_y)
        excess = max(excess, float(np.hypot(x, z)) + r_bed - filter_inner_radius(f, y_local))
        clear = min(clear, y - (bot + r_bed))
    return {"all_finite": finite, "active_count": len(act),
            "max_radial_excess": excess if len(act) else 0.0, "min_floor_clearance": clear if len(act) else 0.0}


def saturated_bed_snapshot(sim) -> dict:
    """readback_saturated_bed_motion_snapshot (physics_tests.rs:838-935):
    compaction is BedExtract.bed.w, saturation BedExtract.extract.w."""
    p = sim.read_bed_particles()[0].astype(np.float64)
    e = sim.read_bed_extracts().astype(np.float64)
    comp, sat = e[:, 3], e[:, 7]
    finite = bool(np.isfinite(p[:, 7]).all() and np.isfinite(comp).all() and np.isfinite(sat).all())
    act = p[:, 7] > INACTIVE
    wet = act & (sat >= 0.65)
    return {"all_finite": finite, "active_count": int(act.sum()), "saturated_count": int(wet.sum()),
            "mean_compression": float(comp[act].sum() / max(int(act.sum()), 1)),
            "saturated_mean_compression": float(comp[wet].sum() / max(int(wet.sum()), 1))}


def bed_states(sim) -> dict:
    """readback_bed_particle_states (physics_tests.rs:695-775): position, mass, saturation."""
    p = sim.read_bed_particles()[0].astype(np.float64)
    return {"pos": p[:, 0:3].copy(), "mass": p[:, 7].copy(), "saturation": sim.read_bed_extracts()[:, 7].astype(np.float64)}


def bed_motion_delta(before: dict, after: dict, active_mass: float, saturated: float, center_radius: float) -> dict:
    """bed_particle_motion_delta_snapshot (physics_tests.rs:777-836): the motion of
    the saturated bed particles within `center_radius` of the axis."""
    n = min(len(before["mass"]), len(after["mass"]))
    b, a = before["pos"][:n], after["pos"][:n]
    finite = bool(np.isfinite(b).all() and np.isfinite(a).all() and np.isfinite(before["mass"][:n]).all()
                  and np.isfinite(after["mass"][:n]).all() and np.isfinite(after["saturation"][:n]).all())
    act = after["mass"][:n] > active_mass
    wet = act & (after["saturation"][:n] >= saturated)
    tracked = wet & (np.hypot(a[:, 0], a[:, 2]) <= center_radius)
    d = np.linalg.norm(a - b, axis=1)[tracked]
    down = np.maximum(b[:, 1] - a[:, 1], 0.0)[tracked]
    k = max(int(tracked.sum()), 1)
    return {"all_finite": finite, "active_count": int(act.sum()), "saturated_count": int(wet.sum()),
            "tracked_count": int(tracked.sum()), "mean_displacement": float(d.sum() / k),
            "max_displacement": float(d.max()) if len(d) else 0.0, "mean_downward_displacement": float(down.sum() / k)}
