"""Seeded synthetic scenes for the step workloads (host-side input generation).

All random draws come from XorShift64 as the reference defines it
(crates/riftphys-core/src/rng.rs:6-39: state = seed | 1, xorshift 12/25/27, output
(x * 2685821657736338717) >> 32, f32 in [0, 1) from the top 23 bits), in the loop order each
scene documents, so a scene is a pure function of its seed.  Seeds: scene k, world w ->
0x5EED0000 + 1000 * k + w.

A scene is a dict of numpy arrays that both the device library (through the C ABI) and the
CPU oracle consume:
  shape  int32[N]      0 sphere, 1 box, 2 capsule
  dims   float32[N,3]  sphere (r,0,0), box (hx,hy,hz), capsule (r,hh,0)
  pose   float32[N,7]  px,py,pz,qx,qy,qz,qw
  vel    float32[N,6]
  dynamic int32[N], density float32[N]
  hf     dict(nx, nz, sx, sz, ox, oz, heights float32[nz*nx]) (row-major, z rows) or None
  pairs  list of (mat_a, mat_b, mu_s, mu_k, restitution, stribeck_v) overrides
  iters  (velocity, position), dt
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1
MAT_DEFAULT = 0
MAT_GRIT = 4


class XorShift64:
    """crates/riftphys-core/src/rng.rs:6-39."""

    def __init__(self, seed):
        self.state = (seed | 1) & MASK64

    def next_u32(self):
        x = self.state
        x ^= x >> 12
        x ^= (x << 25) & MASK64
        x ^= x >> 27
        self.state = x
        return ((x * 2685821657736338717) & MASK64) >> 32

    def next_f32_01(self):
        u = self.next_u32()
        bits = (u >> 9) | 0x3F800000
        return float(np.array([bits], dtype=np.uint32).view(np.float32)[0]) - 1.0

    def uniform(self, lo, hi):
        """lo + (hi - lo) * u, evaluated in f32."""
        u = np.float32(self.next_f32_01())
        return np.float32(np.float32(lo) + np.float32(np.float32(hi) - np.float32(lo)) * u)


def seed_for(scene_k, world=0):
    return 0x5EED0000 + 1000 * scene_k + world


def quat_yaw_pitch(yaw, pitch):
    """q = Ry(yaw) * Rx(pitch) as (x, y, z, w) in f32 (input data, not solver arithmetic)."""
    cy, sy = math.cos(yaw * 0.5), math.sin(yaw * 0.5)
    cp, sp = math.cos(pitch * 0.5), math.sin(pitch * 0.5)
    # Hamilton product (0, sy, 0, cy) * (sp, 0, 0, cp)
    x = cy * sp
    y = sy * cp
    z = -sy * sp
    w = cy * cp
    return np.array([x, y, z, w], dtype=np.float32)


def fbm_heights(rng, n, cell, octaves=4, base_freq=1.0 / 32.0, lacunarity=2.0, gain=0.5, amplitude=4.0):
    """Value-noise fBm on an n x n grid (row-major, z rows), normalised to [0, amplitude]."""
    size = (n - 1) * cell
    xs = np.arange(n, dtype=np.float64) * cell
    total = np.zeros((n, n), dtype=np.float64)
    freq, amp = base_freq, 1.0
    for _ in range(octaves):
        m = int(math.ceil(size * freq)) + 2
        lat = np.array([rng.next_f32_01() for _ in range(m * m)], dtype=np.float64).reshape(m, m)
        u = xs * freq
        i0 = np.floor(u).astype(np.int64)
        f = u - i0
        f = f * f * (3.0 - 2.0 * f)
        # bilinear on the smoothstep weights; rows = z, cols = x
        z0, x0 = i0[:, None], i0[None, :]
        fz, fx = f[:, None], f[None, :]
        v00 = lat[z0, x0]
        v10 = lat[z0, x0 + 1]
        v01 = lat[z0 + 1, x0]
        v11 = lat[z0 + 1, x0 + 1]
        total += amp * ((v00 * (1 - fx) + v10 * fx) * (1 - fz) + (v01 * (1 - fx) + v11 * fx) * fz)
        freq *= lacunarity
        amp *= gain
    total -= total.min()
    peak = total.max()
    if peak > 0:
        total *= amplitude / peak
    return total.astype(np.float32).reshape(-1)


def _empty(n):
    return dict(
        shape=np.zeros(n, np.int32),
        dims=np.zeros((n, 3), np.float32),
        pose=np.zeros((n, 7), np.float32),
        vel=np.zeros((n, 6), np.float32),
        dynamic=np.ones(n, np.int32),
        density=np.full(n, 1000.0, np.float32),
    )


CONFIG_PAIRS = [(MAT_DEFAULT, MAT_DEFAULT, 0.5, 0.5, 0.0, 0.08), (MAT_DEFAULT, MAT_GRIT, 0.5, 0.5, 0.0, 0.08)]


def box_lattice(seed=None, n_side=8, spacing=1.1, half=0.5, hf_n=64):
    """512 boxes (0.5 m half-extent, 1000 kg/m^3) in an 8x8x8 lattice at 1.1 m, jitter U(+-0.05),
    random yaw U(-pi, pi), over a flat 64x64 heightfield of 1 m cells.  Draw order: i, j, k,
    then jx, jy, jz, yaw."""
    rng = XorShift64(seed_for(1) if seed is None else seed)
    n = n_side ** 3
    s = _empty(n)
    c = (n_side - 1) / 2.0
    idx = 0
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                jx = rng.uniform(-0.05, 0.05)
                jy = rng.uniform(-0.05, 0.05)
                jz = rng.uniform(-0.05, 0.05)
                yaw = float(rng.uniform(-math.pi, math.pi))
                p = np.array([32.0 + (i - c) * spacing, 1.05 + spacing * j, 32.0 + (k - c) * spacing], np.float32)
                p = p + np.array([jx, jy, jz], np.float32)
                s["shape"][idx] = 1
                s["dims"][idx] = (half, half, half)
                s["pose"][idx, :3] = p
                s["pose"][idx, 3:] = quat_yaw_pitch(yaw, 0.0)
                idx += 1
    hts = np.zeros(hf_n * hf_n, np.float32)
    s["hf"] = dict(nx=hf_n, nz=hf_n, sx=np.float32(63.0 / (hf_n - 1)), sz=np.float32(63.0 / (hf_n - 1)),
                   ox=0.0, oz=0.0, heights=hts)
    s["pairs"] = list(CONFIG_PAIRS)
    s["iters"] = (10, 20)
    s["dt"] = np.float32(1.0 / 60.0)
    s["name"] = "box_lattice_512"
    return s


def mixed_shapes(n=16384, seed=None, hf_n=256, cell=0.5, extent=100.0, height=30.0, margin=13.75):
    """n bodies: i%3==0 spheres r~U(0.2,0.6); 1 boxes he~U(0.2,0.6)^3; 2 capsules r 0.3,
    hh~U(0.3,0.8); uniform yaw and pitch; x,z ~ U(margin, margin+extent), y ~ U(hmax+1, hmax+1+height),
    over a 4-octave fBm heightfield (amplitude 4 m).  Terrain is drawn first, then per body:
    shape dims, x, y, z, yaw, pitch."""
    rng = XorShift64(seed_for(2) if seed is None else seed)
    hts = fbm_heights(rng, hf_n, cell)
    hmax = float(hts.max())
    s = _empty(n)
    for i in range(n):
        kind = i % 3
        if kind == 0:
            s["shape"][i] = 0
            s["dims"][i] = (rng.uniform(0.2, 0.6), 0.0, 0.0)
        elif kind == 1:
            s["shape"][i] = 1
            s["dims"][i] = (rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6))
        else:
            s["shape"][i] = 2
            s["dims"][i] = (0.3, rng.uniform(0.3, 0.8), 0.0)
        x = rng.uniform(margin, margin + extent)
        y = rng.uniform(hmax + 1.0, hmax + 1.0 + height)
        z = rng.uniform(margin, margin + extent)
        yaw = float(rng.uniform(-math.pi, math.pi))
        pitch = float(rng.uniform(-math.pi, math.pi))
        s["pose"][i, :3] = (x, y, z)
        s["pose"][i, 3:] = quat_yaw_pitch(yaw, pitch)
    s["hf"] = dict(nx=hf_n, nz=hf_n, sx=np.float32(cell), sz=np.float32(cell), ox=0.0, oz=0.0, heights=hts)
    s["pairs"] = list(CONFIG_PAIRS)
    s["iters"] = (10, 20)
    s["dt"] = np.float32(1.0 / 60.0)
    s["name"] = "mixed_%d" % n
    return s


def batch_world(w, seed_base_k=3):
    """One world of the batched workload: N_w = round(exp(U(ln 32, ln 2048))) bodies with the mixed
    shape mix and fBm terrain, volume scaled to keep the 16,384-body density (same 30 m height)."""
    rng = XorShift64(seed_for(seed_base_k, w))
    u = float(rng.next_f32_01())
    nw = int(round(math.exp(math.log(32.0) + u * (math.log(2048.0) - math.log(32.0)))))
    extent = 100.0 * math.sqrt(nw / 16384.0)
    hf_n = min(256, int(math.ceil((extent + 2 * 13.75) / 0.5)) + 1)
    return mixed_shapes(n=nw, seed=seed_for(seed_base_k, w) ^ 0x9E3779B9, hf_n=hf_n, extent=extent)


def small_mixed(n=96, seed=7, hf_n=48):
    """A few dozen mixed bodies over fBm terrain: the quick parity case."""
    return mixed_shapes(n=n, seed=seed, hf_n=hf_n, cell=0.5, extent=14.0, height=6.0, margin=4.0)


def towers_ragdolls(n_towers=669, tower_h=20, n_loose=4, n_ragdolls=200, chain=15, seed=None, hf_n=256, cell=0.5,
                    extent=100.0, margin=13.75, drop=8.0):
    """Box towers and capsule chains over fBm terrain (64 of these worlds, 16,384 bodies each, are
    the at-scale workload): n_towers towers of tower_h boxes (half-extent 0.25 m, 1000 kg/m^3) on a
    grid with jitter U(+-0.2) and one yaw per tower, each box 5 mm above the one below it and the
    lowest 5 mm above the highest terrain sample under its footprint; n_loose boxes above the
    terrain; n_ragdolls chains of `chain` capsules (r 0.1, hh 0.15) lying horizontally at a random
    yaw, end to end, y ~ U(hmax+2, hmax+2+drop), consecutive capsules joined by ball joints at
    their touching ends.  Bodies: towers (bottom to top), loose boxes, chains.  Draw order:
    terrain, per tower (jx, jz, yaw), per loose box (x, y, z, yaw), per chain (x, y, z, yaw)."""
    rng = XorShift64(seed_for(5) if seed is None else seed)
    hts = fbm_heights(rng, hf_n, cell)
    grid = hts.reshape(hf_n, hf_n)
    hmax = float(hts.max())
    n = n_towers * tower_h + n_loose + n_ragdolls * chain
    s = _empty(n)
    g = int(math.ceil(math.sqrt(max(n_towers, 1))))
    spacing = extent / g
    idx = 0

    def ground(x, z, r):
        gx0, gx1 = max(int(math.floor((x - r) / cell)), 0), min(int(math.ceil((x + r) / cell)), hf_n - 1)
        gz0, gz1 = max(int(math.floor((z - r) / cell)), 0), min(int(math.ceil((z + r) / cell)), hf_n - 1)
        return float(grid[gz0:gz1 + 1, gx0:gx1 + 1].max())

    for t in range(n_towers):
        gi, gj = t % g, t // g
        x = float(np.float32(margin + (gi + 0.5) * spacing) + rng.uniform(-0.2, 0.2))
        z = float(np.float32(margin + (gj + 0.5) * spacing) + rng.uniform(-0.2, 0.2))
        yaw = float(rng.uniform(-math.pi, math.pi))
        base = ground(x, z, 0.36) + 0.255
        q = quat_yaw_pitch(yaw, 0.0)
        for k in range(tower_h):
            s["shape"][idx] = 1
            s["dims"][idx] = (0.25, 0.25, 0.25)
            s["pose"][idx, :3] = (x, base + 0.505 * k, z)
            s["pose"][idx, 3:] = q
            idx += 1
    for _ in range(n_loose):
        x = rng.uniform(margin, margin + extent)
        y = rng.uniform(hmax + 2.0, hmax + 2.0 + drop)
        z = rng.uniform(margin, margin + extent)
        yaw = float(rng.uniform(-math.pi, math.pi))
        s["shape"][idx] = 1
        s["dims"][idx] = (0.25, 0.25, 0.25)
        s["pose"][idx, :3] = (x, y, z)
        s["pose"][idx, 3:] = quat_yaw_pitch(yaw, 0.0)
        idx += 1
    joints = []
    half = 0.25  # hh + r: a capsule's end along its local +Y
    for _ in range(n_ragdolls):
        x = rng.uniform(margin, margin + extent)
        y = rng.uniform(hmax + 2.0, hmax + 2.0 + drop)
        z = rng.uniform(margin, margin + extent)
        yaw = float(rng.uniform(-math.pi, math.pi))
        q = quat_yaw_pitch(yaw, math.pi / 2)  # local +Y horizontal
        d = np.array([2.0 * (q[0] * q[1] - q[3] * q[2]), 1.0 - 2.0 * (q[0] ** 2 + q[2] ** 2), 2.0 * (q[1] * q[2] + q[3] * q[0])],
                     np.float64)
        first = idx
        for k in range(chain):
            p = np.array([x, y, z], np.float64) + d * (2.0 * half * k)
            s["shape"][idx] = 2
            s["dims"][idx] = (0.1, 0.15, 0.0)
            s["pose"][idx, :3] = p.astype(np.float32)
            s["pose"][idx, 3:] = q
            idx += 1
        for k in range(chain - 1):
            joints.append((first + k, first + k + 1, (0.0, half, 0.0, 0.0, 0.0, 0.0, 1.0), (0.0, -half, 0.0, 0.0, 0.0, 0.0, 1.0)))
    assert idx == n
    s["joints"] = joints
    s["hf"] = dict(nx=hf_n, nz=hf_n, sx=np.float32(cell), sz=np.float32(cell), ox=0.0, oz=0.0, heights=hts)
    s["pairs"] = list(CONFIG_PAIRS)
    s["iters"] = (10, 2)  # position iterations 2: the workload's two-iteration joint relaxation
    s["dt"] = np.float32(1.0 / 60.0)
    s["name"] = "towers_ragdolls_%d" % n
    return s


def ragdolls_small(seed=11, n_towers=4, tower_h=5, n_ragdolls=6, chain=15, hf_n=48):
    """A few towers and capsule chains over fBm terrain: the quick joint parity case."""
    return towers_ragdolls(n_towers=n_towers, tower_h=tower_h, n_loose=2, n_ragdolls=n_ragdolls, chain=chain, seed=seed,
                           hf_n=hf_n, cell=0.5, extent=14.0, margin=4.0, drop=3.0)


def jointed_bounce_pile(hz=60.0):
    """The W6 pile golden's pyramid with its dropped box, struck by a box on a hinge (a D6 joint
    about z to a static pin at (0, 1.6, 0), released level) and by a box on a rigid distance joint
    swinging in from the side: restitution scenes with jointed bodies (immovable to W6)."""
    s_bodies = [bounce_floor(MAT_CONCRETE)]
    f = np.float32
    for layer, n in enumerate([6, 5, 4, 3, 2]):
        for i in range(n):
            x = (f(i) - (f(n) - f(1.0)) * f(0.5)) * f(0.4)
            s_bodies.append(bounce_body(1, BOUNCE_HALF, (x, f(0.1) + f(0.2) * f(layer), 0.0), 2.0, MAT_ALUMINUM))
    s_bodies.append(bounce_pile_drop())
    k = len(s_bodies)
    s_bodies += [_static_pin((0.0, 1.6, 0.0)), bounce_body(1, BOUNCE_HALF, (0.8, 1.6, 0.0), 2.0, MAT_ALUMINUM),
                 _static_pin((-1.4, 1.7, 0.0)), bounce_body(1, BOUNCE_HALF, (-1.4, 0.9, 0.0), 2.0, MAT_ALUMINUM, vel=(3.0, 0, 0, 0, 0, 0))]
    s = bounce_scene_of(s_bodies, hz=hz, name="jointed_bounce_pile")
    s["d6"] = [dict(a=k, b=k + 1, fa=_iso((0, 0, 0)), fb=_iso((-0.8, 0, 0)), t_mask=7, r_mask=3, hinge=2)]
    s["dist"] = [(k + 2, k + 3, 0.8, 0.0)]
    return s


def d6_ragdolls(seed=11, n_towers=3, tower_h=4, n_ragdolls=4):
    """ragdolls_small with two of every three chain joints made D6 joints: hinges about z limited to
    +-1 rad with a damper, and twist-free swing joints limited to +-0.5 rad about x and y with
    friction on the twist; the rest stay ball joints, all in the same islands as the contacts."""
    s = ragdolls_small(seed=seed, n_towers=n_towers, tower_h=tower_h, n_ragdolls=n_ragdolls)
    balls, d6 = [], []
    for k, (a, b, fa, fb) in enumerate(s["joints"]):
        fa, fb = np.asarray(fa, np.float32), np.asarray(fb, np.float32)
        if k % 3 == 0:
            d6.append(dict(a=a, b=b, fa=fa, fb=fb, t_mask=7, r_mask=3, hinge=2, limits=[(True, 2, -1.0, 1.0)], damping=[(True, 2, 0.5, 0.0)]))
        elif k % 3 == 1:
            d6.append(dict(a=a, b=b, fa=fa, fb=fb, t_mask=7, r_mask=0, hinge=-1, limits=[(True, 0, -0.5, 0.5), (True, 1, -0.5, 0.5)],
                           damping=[(True, 2, 0.0, 0.05)]))
        else:
            balls.append((a, b, fa, fb))
    s["joints"] = balls
    s["d6"] = d6
    return s


# ---- bounce scenes: the restitution cases of crates/riftphys-world/tests/restitution_toi_test.rs
# (static box floor with its top at y = 0, boxes and balls of given mass and material id, sleep
# disabled on the moving bodies, gravity -9.81 or none).  Material ids are the reference's core
# table (materials.rs:11-12).
MAT_ICE, MAT_RUBBER, MAT_CONCRETE, MAT_ALUMINUM = 1, 2, 6, 13


def bounce_body(shape, dims, pos, mass, mat, rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)):
    """One moving body of the restitution tests: a box (dims = half extents) of density
    mass / (8 hx hy hz), or a sphere (dims[0] = r) of density mass / (4/3 pi r^3), each in f32 in
    the tests' order (restitution_toi_test.rs:28-44)."""
    f = np.float32
    d = np.array(dims, f)
    if shape == 1:
        dens = f(mass) / (((f(8.0) * d[0]) * d[1]) * d[2])
    else:
        r = d[0]
        dens = f(mass) / ((((f(4.0) / f(3.0)) * f(math.pi)) * r) * r * r)
    return dict(shape=shape, dims=d, pose=np.array([*pos, *rot], f), vel=np.array(vel, f), dynamic=1, density=f(dens),
                material=mat, sleep_allowed=0)


def bounce_floor(mat):
    """A static 100 x 1 x 100 m box whose top is at y = 0 (restitution_toi_test.rs:15-19)."""
    return dict(shape=1, dims=np.array([50.0, 0.5, 50.0], np.float32), pose=np.array([0, -0.5, 0, 0, 0, 0, 1], np.float32),
                vel=np.zeros(6, np.float32), dynamic=0, density=np.float32(0.0), material=mat, sleep_allowed=1)


def bounce_scene_of(bodies, gravity=True, hz=60.0, name="bounce"):
    n = len(bodies)
    s = dict(shape=np.array([b["shape"] for b in bodies], np.int32), dims=np.zeros((n, 3), np.float32),
             pose=np.array([b["pose"] for b in bodies], np.float32), vel=np.array([b["vel"] for b in bodies], np.float32),
             dynamic=np.array([b["dynamic"] for b in bodies], np.int32),
             density=np.array([b["density"] for b in bodies], np.float32),
             material=np.array([b["material"] for b in bodies], np.int32),
             sleep_allowed=np.array([b["sleep_allowed"] for b in bodies], np.int32))
    for i, b in enumerate(bodies):
        s["dims"][i, :len(b["dims"])] = b["dims"]
    s["gravity"] = (0.0, -9.81, 0.0) if gravity else (0.0, 0.0, 0.0)
    s["hf"] = None
    s["iters"] = (20, 20)  # World::with_capacity's solver iterations (world.rs:425)
    s["dt"] = np.float32(1.0) / np.float32(hz)
    s["name"] = name
    return s


BOUNCE_HALF = (0.2, 0.1, 0.2)


def bounce_pile(hz=60.0):
    """The 20-box pyramid (layers of 6, 5, 4, 3, 2 aluminium boxes on concrete) of the W6 pile
    golden, before its drop; the box dropped after 1 s is bounce_pile_drop()
    (restitution_toi_test.rs:532-553)."""
    f = np.float32
    bodies = [bounce_floor(MAT_CONCRETE)]
    for layer, n in enumerate([6, 5, 4, 3, 2]):
        for i in range(n):
            x = (f(i) - (f(n) - f(1.0)) * f(0.5)) * f(0.4)
            y = f(0.1) + f(0.2) * f(layer)
            bodies.append(bounce_body(1, BOUNCE_HALF, (x, y, 0.0), 2.0, MAT_ALUMINUM))
    return bounce_scene_of(bodies, hz=hz, name="bounce_pile")


def bounce_pile_drop():
    return bounce_body(1, BOUNCE_HALF, (-0.2, 2.0, 0.0), 2.0, MAT_ALUMINUM)


def bounce_stack(hz=60.0):
    """A two-box aluminium stack with a third box 1.2 m above (restitution_toi_test.rs:548-550)."""
    bodies = [bounce_floor(MAT_CONCRETE)] + [bounce_body(1, BOUNCE_HALF, (0.0, y, 0.0), 2.0, MAT_ALUMINUM) for y in (0.1, 0.3, 1.5)]
    return bounce_scene_of(bodies, hz=hz, name="bounce_stack")


def bounce_drone(hz=60.0):
    """The drone-like aluminium box (0.4 x 0.1 x 0.4 m, 1.5 kg) flat from 2 m over concrete
    (restitution_toi_test.rs:612-614)."""
    bodies = [bounce_floor(MAT_CONCRETE), bounce_body(1, (0.2, 0.05, 0.2), (0.0, 2.0, 0.0), 1.5, MAT_ALUMINUM)]
    return bounce_scene_of(bodies, hz=hz, name="bounce_drone")


# Contact-model scenes: a 40 x 1 x 40 m static box whose top is at y = 0 (tilted about z by
# ground_tilt radians around the origin) and one dynamic body of mass 1 kg (box dims = half
# extents, or a sphere of radius dims[0]) that may not sleep, on a pair whose contact model is set
# (compliant_contact_test.rs:10-27, patch_friction_test.rs:16-33).  Friction mu on both
# coefficients; the restitution and Stribeck speed given are unused by compliant rows and by equal
# coefficients.
MAT_STEEL, MAT_PMMA, MAT_WOOD = 3, 5, 9


def contact_model_scene(k, d, mu, shape, dims, pos, vel=(0.0, 0.0, 0.0), ang=(0.0, 0.0, 0.0), hz=1000.0, gravity=True,
                        patch=False, mass=1.0, ground_tilt=0.0, rot=(0.0, 0.0, 0.0, 1.0), rest=0.0, mats=(MAT_STEEL, MAT_CONCRETE),
                        name="contact_model"):
    f = np.float32
    q = (0.0, 0.0, math.sin(ground_tilt / 2.0), math.cos(ground_tilt / 2.0))
    # the ground's centre: the rotation applied to (0, -0.5, 0)
    c, s_ = math.cos(ground_tilt), math.sin(ground_tilt)
    gpos = (0.5 * s_, -0.5 * c, 0.0)
    ground = dict(shape=1, dims=np.array([20.0, 0.5, 20.0], f), pose=np.array([*gpos, *q], f), vel=np.zeros(6, f), dynamic=0,
                  density=f(0.0), material=mats[1], sleep_allowed=1)
    body = bounce_body(shape, dims, pos, mass, mats[0], rot=rot, vel=(*vel, *ang))
    s = bounce_scene_of([ground, body], gravity=gravity, hz=hz, name=name)
    s["pairs"] = [(mats[0], mats[1], f(mu), f(mu), f(rest), f(0.1))]
    s["contact_models"] = [(mats[0], mats[1], f(k), f(d), bool(patch))]
    return s


# D6 joint scenes (hinge_angle_test.rs, joint_damping_test.rs, joint_energy_test.rs,
# joint_bits_probe.rs): static pivots carry a 1 cm sphere collider; rods are 1 m x 5 cm x 5 cm
# boxes of density 1000 (2.5 kg) lying along +x.  A "d6" entry names bodies by scene index, its
# frames {p, q} in each body's space, the constrained axes (t_mask, r_mask), a stored hinge axis,
# and per axis (angular flag, axis, ...) limits, drives, damping and friction, compliance.
ROD_HALF = (0.5, 0.025, 0.025)


def _static_pin(pos):
    return dict(shape=0, dims=np.array([0.01, 0.0, 0.0], np.float32), pose=np.array([*pos, 0, 0, 0, 1], np.float32),
                vel=np.zeros(6, np.float32), dynamic=0, density=np.float32(0.0), material=0, sleep_allowed=1)


def _box_body(half, pos, density, vel=(0, 0, 0, 0, 0, 0), sleep=1):
    return dict(shape=1, dims=np.array(half, np.float32), pose=np.array([*pos, 0, 0, 0, 1], np.float32), vel=np.array(vel, np.float32),
                dynamic=1, density=np.float32(density), material=0, sleep_allowed=sleep)


def _iso(p):
    return np.array([*p, 0.0, 0.0, 0.0, 1.0], np.float32)


def _joint_scene(bodies, joints, gravity=True, hz=240.0, name="d6"):
    s = bounce_scene_of(bodies, gravity=gravity, hz=hz, name=name)
    s["d6"] = joints
    return s


def hinge_rod(limit=None, hz=240.0, drive=None, damping=None, compliance=None, gravity=True):
    """A rod hinged at its end about z to a static pivot at (0, 10, 0), lying along +x."""
    j = dict(a=0, b=1, fa=_iso((0, 0, 0)), fb=_iso((-0.5, 0, 0)), t_mask=7, r_mask=3, hinge=2)
    if limit is not None:
        j["limits"] = [(True, 2, limit[0], limit[1])]
    if drive is not None:
        j["drives"] = [(True, 2, *drive)]
    if damping is not None:
        j["damping"] = [(True, 2, *damping)]
    if compliance is not None:  # a sprung hinge: the twist row enabled, compliant
        j["compliance"] = [(True, 2, compliance)]
        j["r_mask"] = 7
    return _joint_scene([_static_pin((0, 10, 0)), _box_body(ROD_HALF, (0.5, 10, 0), 1000.0, sleep=0)], [j], gravity, hz, "hinge_rod")


SPIN_HALF, SPIN_MASS = (0.5, 0.1, 0.1), 2.0


def hinge_spinner(w0, damping=0.0, friction=0.0, hz=60.0):
    """A 2 kg bar hinged about z through its centre to a static anchor, spinning at w0, no gravity."""
    d = SPIN_MASS / (8.0 * SPIN_HALF[0] * SPIN_HALF[1] * SPIN_HALF[2])
    j = dict(a=0, b=1, fa=_iso((0, 0, 0)), fb=_iso((0, 0, 0)), t_mask=7, r_mask=3, hinge=2, damping=[(True, 2, damping, friction)])
    bar = _box_body(SPIN_HALF, (0, 0, 0), d, vel=(0, 0, 0, 0, 0, w0), sleep=0)
    return _joint_scene([_static_pin((0, 0, 0)), bar], [j], False, hz, "hinge_spinner")


def cube_pendulum(hz=62.5):
    """An 80 kg, 2 m cube on a 5 m arm hinged about z at (0, 20, 0), released horizontal."""
    j = dict(a=0, b=1, fa=_iso((0, 0, 0)), fb=_iso((-5.0, 0, 0)), t_mask=7, r_mask=3, hinge=2)
    return _joint_scene([_static_pin((0, 20, 0)), _box_body((1.0, 1.0, 1.0), (5.0, 20, 0), 10.0, sleep=0)], [j], True, hz, "cube_pendulum")


def hinge_chain(hz=240.0):
    """Four rods hinged end to end about z from a static pivot: the first limited to +-1 rad, the
    second a velocity motor (2 rad/s, at most 0.05 N s a step), the rest free
    (joint_bits_probe.rs)."""
    bodies = [_static_pin((0, 10, 0))] + [_box_body(ROD_HALF, (0.5 + k, 10, 0), 1000.0, sleep=0) for k in range(4)]
    joints = []
    for k in range(4):
        j = dict(a=k, b=k + 1, fa=_iso((0, 0, 0) if k == 0 else (0.5, 0, 0)), fb=_iso((-0.5, 0, 0)), t_mask=7, r_mask=3, hinge=2)
        if k == 0:
            j["limits"] = [(True, 2, -1.0, 1.0)]
        if k == 1:
            j["drives"] = [(True, 2, 2.0, 0.0, 0.0, 0.05, True)]
        joints.append(j)
    return _joint_scene(bodies, joints, True, hz, "hinge_chain")
