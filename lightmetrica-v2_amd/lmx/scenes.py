"""Deterministic procedural scenes and ray sets for the benchmark configs.

All generators are pure numpy and seeded; the same call yields the same
arrays on every machine.  Triangles are returned in enumeration order
(primitive order x face order), the order in which the reference numbers its
triangles (src/liblightmetrica/accel/accel_qbvh.cpp:164-189).
"""
import numpy as np

FLT_MAX = np.float32(np.finfo(np.float32).max)
T_MIN = np.float32(1e-4)      # Math::EpsIsect(), include/lightmetrica/math.h:1668

_ICO_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array([
    [-1, _ICO_T, 0], [1, _ICO_T, 0], [-1, -_ICO_T, 0], [1, -_ICO_T, 0],
    [0, -1, _ICO_T], [0, 1, _ICO_T], [0, -1, -_ICO_T], [0, 1, -_ICO_T],
    [_ICO_T, 0, -1], [_ICO_T, 0, 1], [-_ICO_T, 0, -1], [-_ICO_T, 0, 1]], dtype=np.float64)
_ICO_F = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def icosphere(subdivisions):
    """Unit icosphere: the icosahedron split `subdivisions` times
    (20 * 4**subdivisions triangles, outward counter-clockwise winding).
    Returns float64 vertices (V, 3) and int64 faces (F, 3)."""
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(subdivisions):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a0 = uniq // nv
        a1 = uniq % nv
        mid = v[a0] + v[a1]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], axis=0)
        nf = f.shape[0]
        m01 = nv + inv[:nf]
        m12 = nv + inv[nf:2 * nf]
        m20 = nv + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], axis=1),
            np.stack([f[:, 1], m12, m01], axis=1),
            np.stack([f[:, 2], m20, m12], axis=1),
            np.stack([m01, m12, m20], axis=1)], axis=0)
    return v, f


def _hash3(ix, iy, iz, seed):
    """32-bit integer hash of lattice coordinates -> uniform in [-1, 1]."""
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) \
        ^ (iz.astype(np.uint32) * np.uint32(0xCB1AB31F)) ^ np.uint32(seed * 0x9E3779B9 & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise(p, seed, octaves=3, base_freq=4.0):
    """Value noise: hashed lattice values, smoothstep-trilinear interpolation,
    `octaves` octaves of doubling frequency and halving amplitude, normalised
    to [-1, 1]."""
    total = np.zeros(p.shape[0])
    amp_sum = 0.0
    amp = 1.0
    freq = base_freq
    with np.errstate(over="ignore"):
        for o in range(octaves):
            q = p * freq
            i0 = np.floor(q).astype(np.int64)
            fr = q - i0
            w = fr * fr * (3.0 - 2.0 * fr)
            acc = np.zeros(p.shape[0])
            for dx in (0, 1):
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                for dy in (0, 1):
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    for dz in (0, 1):
                        wz = w[:, 2] if dz else 1.0 - w[:, 2]
                        acc += wx * wy * wz * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + 7919 * o)
            total += amp * acc
            amp_sum += amp
            amp *= 0.5
            freq *= 2.0
    return total / amp_sum


def displaced_icosphere(subdivisions, amplitude=0.1, seed=1, center=(0.0, 0.0, 0.0), radius=1.0):
    """Icosphere with vertices displaced radially by 3-octave value noise:
    r = radius * (1 + amplitude * noise(unit position)).  float32 verts, uint32 tris."""
    v, f = icosphere(subdivisions)
    n = value_noise(v, seed)
    v = v * (radius * (1.0 + amplitude * n))[:, None] + np.asarray(center, dtype=np.float64)[None, :]
    return v.astype(np.float32), f.astype(np.uint32)


def make_rays(o, d, tmin=T_MIN, tmax=FLT_MAX):
    """Pack origins and directions into (n, 8) float32 ray records."""
    n = o.shape[0]
    r = np.empty((n, 8), dtype=np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def pinhole_rays(width, height, eye, center, up, fov_deg, tmin=T_MIN):
    """Primary rays through pixel centres of a pinhole camera (vertical fov),
    row 0 at the bottom as in the reference's raster convention
    (src/liblightmetrica/asset/sensor/sensor_pinhole.cpp:79-90)."""
    eye = np.asarray(eye, dtype=np.float64)
    vz = eye - np.asarray(center, dtype=np.float64)
    vz /= np.linalg.norm(vz)
    vx = np.cross(np.asarray(up, dtype=np.float64), vz)
    vx /= np.linalg.norm(vx)
    vy = np.cross(vz, vx)
    tf = np.tan(np.radians(fov_deg) * 0.5)
    aspect = width / height
    xs = (np.arange(width) + 0.5) / width * 2.0 - 1.0
    ys = (np.arange(height) + 0.5) / height * 2.0 - 1.0
    X, Y = np.meshgrid(xs, ys)
    de = np.stack([aspect * tf * X.ravel(), tf * Y.ravel(), -np.ones(X.size)], axis=1)
    de /= np.linalg.norm(de, axis=1, keepdims=True)
    d = de[:, 0:1] * vx + de[:, 1:2] * vy + de[:, 2:3] * vz
    o = np.broadcast_to(eye, d.shape)
    return make_rays(o.astype(np.float32), d.astype(np.float32), tmin)


def incoherent_rays(n, lo, hi, seed, tmin=T_MIN):
    """Origins uniform in the box [lo, hi], directions uniform on the sphere
    by the reference's mapping z = 1 - 2 u0, phi = 2 pi u1
    (include/lightmetrica/sampler.h:79-85)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    o = lo + (hi - lo) * rng.random((n, 3))
    u = rng.random((n, 2))
    z = 1.0 - 2.0 * u[:, 0]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * u[:, 1]
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    return make_rays(o.astype(np.float32), d.astype(np.float32), tmin)


# BVH traversal microbenchmark: 1,310,720-triangle displaced icosphere,
# 2048x2048 coherent primary rays + as many incoherent rays.
TRAVERSAL_SUBDIV = 8
TRAVERSAL_SEED = 42
TRAVERSAL_CAMERA = dict(eye=(0.0, 0.6, 3.2), center=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=40.0)


def traversal_scene(subdiv=TRAVERSAL_SUBDIV):
    return displaced_icosphere(subdiv, amplitude=0.1, seed=TRAVERSAL_SEED)


def traversal_rays(verts, res=2048, n_incoherent=None, seed=TRAVERSAL_SEED + 1):
    """(coherent, incoherent) ray sets of the traversal benchmark."""
    coh = pinhole_rays(res, res, **TRAVERSAL_CAMERA)
    if n_incoherent is None:
        n_incoherent = res * res
    lo = verts.min(axis=0)
    hi = verts.max(axis=0)
    inc = incoherent_rays(n_incoherent, lo, hi, seed)
    return coh, inc


# ---------------------------------------------------------------------------
# Renderer scenes.  A scene is a dict:
#   verts (V,3) f32, tris (F,3) u32, normals (V,3) f32 or None,
#   tri_mat (F,) u32, materials: list of dicts {type, R, exponent, Le},
#   lights: list of (first triangle, count) -- one entry per light primitive,
#   camera: dict(eye, center, up, fov_deg, width, height).
MAT_NULL = 0
MAT_DIFFUSE = 1
MAT_PHONG = 2
MAT_GGX = 3


def _quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); the face normal is
    cross(p1-p0, p2-p0), the winding the reference's gn uses."""
    return [p0, p1, p2, p3], [[0, 1, 2], [0, 2, 3]]


def _box(center, size, angle_deg):
    """Axis-aligned box of `size` rotated about +y, faces outward."""
    cx, cy, cz = center
    sx, sy, sz = (0.5 * s for s in size)
    a = np.radians(angle_deg)
    ca, sa = np.cos(a), np.sin(a)

    def P(x, y, z):
        return [cx + ca * x + sa * z, cy + y, cz - sa * x + ca * z]
    c = [P(-sx, -sy, -sz), P(sx, -sy, -sz), P(sx, sy, -sz), P(-sx, sy, -sz),
         P(-sx, -sy, sz), P(sx, -sy, sz), P(sx, sy, sz), P(-sx, sy, sz)]
    faces = [(4, 5, 6, 7), (1, 0, 3, 2), (0, 4, 7, 3), (5, 1, 2, 6), (7, 6, 2, 3), (0, 1, 5, 4)]
    v, t = [], []
    for f in faces:
        base = len(v)
        v += [c[i] for i in f]
        t += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    return v, t


def cornell_box(width=256, height=256, phong_boxes=None, ggx_boxes=None):
    """Procedural Cornell box: 5 wall quads (inward) + 2 boxes rotated
    +-17 degrees (34 triangles) and a 0.25 x 0.25 emissive ceiling quad facing
    down with Le = (17, 12, 4); Lambertian 0.73 grey / (0.65, 0.05, 0.05) red
    / (0.12, 0.45, 0.15) green; pinhole camera with a 40 degree vertical fov.
    The room is [-1, 1]^3, open towards +z."""
    mats = [dict(type=MAT_DIFFUSE, R=(0.73, 0.73, 0.73), exponent=0.0, Le=(0, 0, 0)),
            dict(type=MAT_DIFFUSE, R=(0.65, 0.05, 0.05), exponent=0.0, Le=(0, 0, 0)),
            dict(type=MAT_DIFFUSE, R=(0.12, 0.45, 0.15), exponent=0.0, Le=(0, 0, 0)),
            dict(type=MAT_NULL, R=(0, 0, 0), exponent=0.0, Le=(17.0, 12.0, 4.0))]
    V, T, M = [], [], []

    def add(vs, ts, m):
        base = len(V)
        V.extend(vs)
        T.extend([[base + i for i in t] for t in ts])
        M.extend([m] * len(ts))
    add(*_quad([-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]), 0)     # floor, normal +y
    add(*_quad([-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]), 0)         # ceiling, normal -y
    add(*_quad([-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]), 0)     # back, normal +z
    add(*_quad([-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]), 1)     # left (red), normal +x
    add(*_quad([1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]), 2)         # right (green), normal -x
    box_mat = 0
    if phong_boxes is not None:        # variant: glossy boxes, normalized Phong of this exponent
        box_mat = len(mats)
        mats.append(dict(type=MAT_PHONG, R=(0.8, 0.8, 0.8), exponent=float(phong_boxes), Le=(0, 0, 0)))
    if ggx_boxes is not None:          # variant: metal boxes, bsdf::cooktorrance of this roughness
        box_mat = len(mats)
        mats.append(dict(type=MAT_GGX, R=(0.9, 0.8, 0.7), exponent=0.0, Le=(0, 0, 0), roughness=float(ggx_boxes)))
    add(*_box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), 17.0), box_mat)        # tall box
    add(*_box((0.35, -0.7, 0.3), (0.6, 0.6, 0.6), -17.0), box_mat)         # short box
    first_light = len(T)
    L = 0.125
    add(*_quad([-L, 0.999, -L], [L, 0.999, -L], [L, 0.999, L], [-L, 0.999, L]), 3)   # light, normal -y
    verts = np.array(V, np.float32)
    tris = np.array(T, np.uint32)
    return dict(verts=verts, tris=tris, normals=None, tri_mat=np.array(M, np.uint32), materials=mats,
                lights=[(first_light, 2)],
                camera=dict(eye=(0.0, 0.0, 3.8), center=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=40.0,
                            width=width, height=height))


def camera_frame(cam):
    """Pinhole camera basis as the reference builds it from a lookat
    transform (scene3.cpp:191-224): vz = normalize(eye - center),
    vx = normalize(cross(up, vz)), vy = cross(vz, vx); the camera looks along
    -vz."""
    eye = np.asarray(cam["eye"], np.float64)
    vz = eye - np.asarray(cam["center"], np.float64)
    vz /= np.linalg.norm(vz)
    vx = np.cross(np.asarray(cam["up"], np.float64), vz)
    vx /= np.linalg.norm(vx)
    vy = np.cross(vz, vx)
    return eye.astype(np.float32), vx.astype(np.float32), vy.astype(np.float32), vz.astype(np.float32)


def furnace_sphere(albedo=0.5, Le=1.0, subdiv=3, width=32, height=32):
    """Closed sphere seen from inside, every triangle Lambertian with albedo
    rho and emission Le (inward winding).  Every camera path sees radiance
    Le * sum_{j < m-1} rho^j with m = max_num_vertices (Le / (1 - rho) when
    unbounded): a closed form for the whole sampler."""
    v, f = icosphere(subdiv)
    f = f[:, [0, 2, 1]]                       # inward normals
    mats = [dict(type=MAT_DIFFUSE, R=(albedo,) * 3, exponent=0.0, Le=(Le,) * 3)]
    return dict(verts=v.astype(np.float32) * 2.0, tris=f.astype(np.uint32), normals=None,
                tri_mat=np.zeros(f.shape[0], np.uint32), materials=mats, lights=[(0, f.shape[0])],
                camera=dict(eye=(0.0, 0.0, 0.3), center=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_deg=60.0,
                            width=width, height=height))


def vertex_normals(v, f):
    """Area-weighted vertex normals (float32, unit length)."""
    v64 = v.astype(np.float64)
    fn = np.cross(v64[f[:, 1]] - v64[f[:, 0]], v64[f[:, 2]] - v64[f[:, 0]])
    n = np.zeros_like(v64)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)


def sphere_grid(width=1024, height=1024, grid=16, subdiv=5, seed=7):
    """Wavefront path-tracer benchmark scene: a grid x grid array of displaced
    icospheres (subdiv 5: 20,480 triangles each; 5,242,880 for the 16 x 16
    grid) on a ground plane, smooth vertex normals; per-sphere material from
    `seed`: 70% Lambertian with albedo U[0.1, 0.9] per channel, 30% normalized
    Phong with exponent log-uniform in [10, 1000]; two emissive quads facing
    down."""
    rng = np.random.default_rng(seed)
    base_v, base_f = icosphere(subdiv)
    nv, nf = base_v.shape[0], base_f.shape[0]
    spacing, radius = 1.0, 0.38
    half = (grid - 1) * spacing * 0.5
    V = np.empty((grid * grid * nv + 12, 3), np.float32)
    F = np.empty((grid * grid * nf + 6, 3), np.uint32)
    M = np.empty(grid * grid * nf + 6, np.uint32)
    mats = []
    for gi in range(grid * grid):
        gx, gz = gi % grid, gi // grid
        c = np.array([gx * spacing - half, radius, gz * spacing - half])
        noise = value_noise(base_v, seed * 1000 + gi)
        v = base_v * (radius * (1.0 + 0.1 * noise))[:, None] + c
        V[gi * nv:(gi + 1) * nv] = v
        F[gi * nf:(gi + 1) * nf] = base_f + gi * nv
        M[gi * nf:(gi + 1) * nf] = len(mats)
        R = tuple(float(x) for x in rng.uniform(0.1, 0.9, 3))
        if rng.random() < 0.7:
            mats.append(dict(type=MAT_DIFFUSE, R=R, exponent=0.0, Le=(0, 0, 0)))
        else:
            e = float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
            mats.append(dict(type=MAT_PHONG, R=R, exponent=e, Le=(0, 0, 0)))
    normals = np.empty_like(V)
    for gi in range(grid * grid):
        normals[gi * nv:(gi + 1) * nv] = vertex_normals(V[gi * nv:(gi + 1) * nv], base_f)
    o = grid * grid * nv
    of = grid * grid * nf
    G = half + 6.0
    # ground (Lambertian 0.5), normal +y
    V[o:o + 4] = [[-G, 0, G], [G, 0, G], [G, 0, -G], [-G, 0, -G]]
    normals[o:o + 4] = [0, 1, 0]
    F[of:of + 2] = [[o, o + 1, o + 2], [o, o + 2, o + 3]]
    M[of:of + 2] = len(mats)
    mats.append(dict(type=MAT_DIFFUSE, R=(0.5, 0.5, 0.5), exponent=0.0, Le=(0, 0, 0)))
    # two area lights facing down
    lights = []
    for k, (cx, cz, le) in enumerate(((-4.0, -2.0, (14.0, 12.0, 10.0)), (4.5, 3.0, (8.0, 10.0, 14.0)))):
        vo = o + 4 + 4 * k
        L = 2.0
        V[vo:vo + 4] = [[cx - L, 7.0, cz - L], [cx + L, 7.0, cz - L], [cx + L, 7.0, cz + L], [cx - L, 7.0, cz + L]]
        normals[vo:vo + 4] = [0, -1, 0]
        ft = of + 2 + 2 * k
        F[ft:ft + 2] = [[vo, vo + 1, vo + 2], [vo, vo + 2, vo + 3]]
        M[ft:ft + 2] = len(mats)
        mats.append(dict(type=MAT_NULL, R=(0, 0, 0), exponent=0.0, Le=le))
        lights.append((ft, 2))
    return dict(verts=V, tris=F, normals=normals, tri_mat=M, materials=mats, lights=lights,
                camera=dict(eye=(0.0, 7.5, 15.5), center=(0.0, 0.0, 0.5), up=(0.0, 1.0, 0.0), fov_deg=45.0,
                            width=width, height=height))


def occluded_room(width=1024, height=1024, n_spheres=96, subdiv=5, seed=11):
    """Path-guiding benchmark scene: a room of 12 quads whose only emitter
    sits in a side compartment behind a full-height dividing wall with a
    0.1-wide gap, so light reaches the main room only through the gap; the
    main room holds n_spheres displaced icospheres (subdiv 5: 20,480
    triangles each, 1,966,080 for 96) of which 40% are normalized Phong with
    exponent 200 and the rest Lambertian.  Quads: floor, ceiling, 4 walls
    (inward), the divider's two pieces with a face on each side (4), the
    emitter (facing the gap) and a shelf in the compartment."""
    rng = np.random.default_rng(seed)
    X0, X1, XD, Y1, Z = -1.5, 1.5, 0.9, 2.0, 1.5
    gap = 0.05
    mats = [dict(type=MAT_DIFFUSE, R=(0.7, 0.7, 0.7), exponent=0.0, Le=(0, 0, 0)),     # 0 walls
            dict(type=MAT_DIFFUSE, R=(0.6, 0.55, 0.45), exponent=0.0, Le=(0, 0, 0)),   # 1 divider
            dict(type=MAT_NULL, R=(0, 0, 0), exponent=0.0, Le=(60.0, 52.0, 44.0))]     # 2 emitter
    V, T, M, N = [], [], [], []

    def add(vs, ts, m):
        base = len(V)
        V.extend(vs)
        T.extend([[base + i for i in t] for t in ts])
        M.extend([m] * len(ts))
        n = np.cross(np.subtract(vs[1], vs[0]), np.subtract(vs[2], vs[0]))
        N.extend([list(n / np.linalg.norm(n))] * len(vs))
    add(*_quad([X0, 0, Z], [X1, 0, Z], [X1, 0, -Z], [X0, 0, -Z]), 0)           # floor, +y
    add(*_quad([X0, Y1, -Z], [X1, Y1, -Z], [X1, Y1, Z], [X0, Y1, Z]), 0)       # ceiling, -y
    add(*_quad([X0, 0, -Z], [X1, 0, -Z], [X1, Y1, -Z], [X0, Y1, -Z]), 0)       # back, +z
    add(*_quad([X1, 0, Z], [X0, 0, Z], [X0, Y1, Z], [X1, Y1, Z]), 0)           # front, -z
    add(*_quad([X0, 0, Z], [X0, 0, -Z], [X0, Y1, -Z], [X0, Y1, Z]), 0)         # left, +x
    add(*_quad([X1, 0, -Z], [X1, 0, Z], [X1, Y1, Z], [X1, Y1, -Z]), 0)         # right, -x
    for z0, z1 in ((-Z, -gap), (gap, Z)):                                      # divider pieces
        add(*_quad([XD, 0, z0], [XD, 0, z1], [XD, Y1, z1], [XD, Y1, z0]), 1)   # main-room face, -x
        add(*_quad([XD, 0, z1], [XD, 0, z0], [XD, Y1, z0], [XD, Y1, z1]), 1)   # compartment face, +x
    first_light = len(T)
    xl = X1 - 0.001
    add(*_quad([xl, 0.8, -0.25], [xl, 0.8, 0.25], [xl, 1.3, 0.25], [xl, 1.3, -0.25]), 2)   # emitter, -x
    add(*_quad([XD, 0.5, Z], [X1, 0.5, Z], [X1, 0.5, -Z], [XD, 0.5, -Z]), 0)   # shelf, +y
    base_v, base_f = icosphere(subdiv)
    nv = base_v.shape[0]
    Vs, Fs, Ms, Ns = [np.array(V, np.float32)], [np.array(T, np.uint32)], [np.array(M, np.uint32)], [np.array(N, np.float32)]
    off = len(V)
    for k in range(n_spheres):
        r = rng.uniform(0.08, 0.16)
        c = np.array([rng.uniform(X0 + 0.2, XD - 0.2), rng.uniform(r, 1.4), rng.uniform(-Z + 0.2, Z - 0.2)])
        noise = value_noise(base_v, seed * 1000 + k)
        v = (base_v * (r * (1.0 + 0.1 * noise))[:, None] + c).astype(np.float32)
        Vs.append(v)
        Fs.append((base_f + off).astype(np.uint32))
        Ns.append(vertex_normals(v, base_f))
        if rng.random() < 0.4:
            mats.append(dict(type=MAT_PHONG, R=tuple(float(x) for x in rng.uniform(0.5, 0.9, 3)),
                             exponent=200.0, Le=(0, 0, 0)))
        else:
            mats.append(dict(type=MAT_DIFFUSE, R=tuple(float(x) for x in rng.uniform(0.2, 0.8, 3)),
                             exponent=0.0, Le=(0, 0, 0)))
        Ms.append(np.full(base_f.shape[0], len(mats) - 1, np.uint32))
        off += nv
    return dict(verts=np.concatenate(Vs), tris=np.concatenate(Fs), normals=np.concatenate(Ns),
                tri_mat=np.concatenate(Ms), materials=mats, lights=[(first_light, 2)],
                camera=dict(eye=(-1.35, 1.1, 1.2), center=(0.9, 0.55, -0.3), up=(0.0, 1.0, 0.0), fov_deg=62.0,
                            width=width, height=height))
