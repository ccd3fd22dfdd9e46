"""Scene descriptions of the benchmark workloads, in the dict form that
rumi.Scene (and the oracle) consume.

A scene dict has "materials" (list of dicts with "kind" in matte, glossy,
specular, emissive, dielectric, subsurface and the reference constructor
parameters), "shapes" (tri / sphere / rect / mesh with a "material" index and
an optional "light" flag), "camera" (position, row-major basis, fov_deg,
y_ratio) and an optional "medium".

The Cornell box is the unit-scale box x in [-0.5, 0.5], y in [0, 1],
z in [-0.5, 0.5], open towards the camera at +z.
"""
from __future__ import annotations

import numpy as np

WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
LIGHT_LE = (17.0, 12.0, 4.0)

# RotationY(pi): camera +z looks down world -z, camera +x maps to world -x.
BASIS_LOOK_MINUS_Z = ((-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, -1.0))


def _quad(a, b, c, d, material, normal):
    """Two triangles a-b-c and a-c-d with a constant vertex normal."""
    n = [normal] * 3
    return [dict(kind="tri", v=[a, b, c], n=n, material=material),
            dict(kind="tri", v=[a, c, d], n=n, material=material)]


def cornell_box(width: int = 512, height: int = 512) -> dict:
    """Five Lambertian walls as 10 triangles, a 0.13-unit quad light of radiance
    (17, 12, 4), a diffuse sphere (r = 0.165) and a glass sphere (r = 0.1,
    IOR 1.5); pinhole camera with a 40 degree field of view."""
    mats = [
        dict(kind="matte", kd=1.0, color=WHITE),           # 0
        dict(kind="matte", kd=1.0, color=RED),             # 1
        dict(kind="matte", kd=1.0, color=GREEN),           # 2
        dict(kind="emissive", kd=1.0, color=LIGHT_LE),     # 3
        dict(kind="dielectric", eta_in=1.5, eta_out=1.0, absorption=(0, 0, 0)),  # 4
    ]
    x0, x1, y0, y1, z0, z1 = -0.5, 0.5, 0.0, 1.0, -0.5, 0.5
    shapes = []
    shapes += _quad((x0, y0, z1), (x1, y0, z1), (x1, y0, z0), (x0, y0, z0), 0, (0, 1, 0))   # floor
    shapes += _quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1), 0, (0, -1, 0))  # ceiling
    shapes += _quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), 0, (0, 0, 1))   # back
    shapes += _quad((x0, y0, z1), (x0, y0, z0), (x0, y1, z0), (x0, y1, z1), 1, (1, 0, 0))   # left
    shapes += _quad((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0), 2, (-1, 0, 0))  # right
    s = 0.13
    shapes.append(dict(kind="rect", p=(-s / 2, 0.999, -s / 2), l=(s, 0, 0), w=(0, 0, s),
                       material=3, light=True))
    shapes.append(dict(kind="sphere", center=(-0.2, 0.165, -0.15), radius=0.165, material=0))
    shapes.append(dict(kind="sphere", center=(0.2, 0.1, 0.15), radius=0.1, material=4))
    cam = dict(position=(0.0, 0.5, 1.9), basis=BASIS_LOOK_MINUS_Z, fov_deg=40.0,
               y_ratio=height / width)
    return dict(materials=mats, shapes=shapes, camera=cam, medium=None)


def cornell_media(width: int = 512, height: int = 512, sigma_a=(0.05, 0.2, 0.4),
                  sigma_s=(1.0, 1.0, 1.0), g=0.7, sss_sigma_a=0.01, sss_sigma_s=2.0) -> dict:
    """The Cornell box with a homogeneous-medium sphere (r = 0.3, HG phase)
    in place of the diffuse sphere and a subsurface sphere (IOR 1.33) in
    place of the glass one."""
    sc = cornell_box(width, height)
    sc["shapes"] = [s for s in sc["shapes"] if s["kind"] != "sphere"]
    sc["materials"].append(dict(kind="subsurface", eta_in=1.33, eta_out=1.0,
                                sigma_s=(sss_sigma_s,) * 3, sigma_a=(sss_sigma_a,) * 3,
                                density=1.0))
    sc["shapes"].append(dict(kind="sphere", center=(0.25, 0.12, 0.15), radius=0.12,
                             material=len(sc["materials"]) - 1))
    sc["medium"] = dict(bound="sphere", center=(-0.18, 0.32, -0.12), radius=0.3,
                        sigma_s=tuple(sigma_s), sigma_a=tuple(sigma_a), g=g)
    return sc


def icosphere_mesh(freq: int, seed: int, amplitude: float, center, scale: float):
    """A seeded noisy icosphere (rumi_make_icosphere) moved to `center` and
    scaled; returns (xyz, idx)."""
    import importlib
    rumi = importlib.import_module("rumi-renderer_amd")
    xyz, idx = rumi.make_icosphere(freq, seed, amplitude)
    xyz = (xyz * np.float32(scale) + np.asarray(center, np.float32)).astype(np.float32)
    return xyz, idx


def cornell_bdpt(width: int = 1024, height: int = 1024, freq: int = 100,
                 roughness: float = 0.2) -> dict:
    """The Cornell box with a seeded noisy icosphere (20 freq^2 triangles,
    200,000 at freq 100, PCG32 seed 3, radius 1 +- 0.05) in GGX material of
    roughness 0.2 in place of the diffuse sphere; the glass sphere stays."""
    sc = cornell_box(width, height)
    sc["shapes"] = [s for s in sc["shapes"] if not (s["kind"] == "sphere" and s["material"] == 0)]
    sc["materials"].append(dict(kind="ggx", color=(0.9, 0.9, 0.9), roughness=roughness))
    xyz, idx = icosphere_mesh(freq, 3, 0.05, center=(-0.2, 0.17, -0.15), scale=0.16)
    sc["shapes"].append(dict(kind="mesh", xyz=xyz, idx=idx, material=len(sc["materials"]) - 1))
    return sc


def diffuse_box(width: int = 64, height: int = 64) -> dict:
    """The Cornell box with both spheres Lambertian (no dirac vertices)."""
    sc = cornell_box(width, height)
    sc["materials"][4] = dict(kind="matte", kd=1.0, color=(0.5, 0.5, 0.5))
    return sc


def _pcg_uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from PCG32(seed, stream): (x >> 8) * 2^-24."""
    import importlib
    rumi = importlib.import_module("rumi-renderer_amd")
    return (rumi.pcg32_sequence(seed, stream, n) >> 8).astype(np.float64) * (1.0 / (1 << 24))


def batch_params(n_scenes: int = 256, seed: int = 4):
    """Per-scene (triangle target T, icosphere frequency, image side) of the
    batched workload: T log-uniform in [1e3, 2e5], frequency round(sqrt(T/20)),
    side log-uniform in [128, 512]; uniforms from PCG32(seed, stream = scene)."""
    out = []
    for i in range(n_scenes):
        u = _pcg_uniform(seed, i, 2)
        t = float(np.exp(np.log(1e3) + u[0] * (np.log(2e5) - np.log(1e3))))
        freq = max(1, int(round(np.sqrt(t / 20.0))))
        side = int(round(np.exp(np.log(128.0) + u[1] * np.log(4.0))))
        out.append((t, freq, side))
    return out


def batch_scene(i: int, freq: int, side: int, seed: int = 4, tex_size: int = 128,
                tex_scale: float = 1.0) -> dict:
    """Scene i of the batched workload: a noisy icosphere (PCG seed 1 + i,
    amplitude 0.05) of 20 freq^2 triangles in Matte material with a
    tex_size^2 RGB albedo texture (uniform in [0.1, 0.9] from PCG32(seed + i),
    times tex_scale), mapped by per-vertex spherical uv; a Lambertian floor;
    a 2 x 2 area light of radiance 6 above; a 40 degree pinhole camera."""
    xyz, idx = icosphere_mesh(freq, 1 + i, 0.05, center=(0.0, 0.0, 0.0), scale=1.0)
    p = xyz.astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    uv = np.stack([np.arctan2(p[:, 0], p[:, 2]) / (2 * np.pi) + 0.5,
                   1.0 - np.arccos(np.clip(p[:, 1] / r, -1, 1)) / np.pi], axis=1).astype(np.float32)
    tex = (0.1 + 0.8 * _pcg_uniform(seed + i, 0, tex_size * tex_size * 3)).astype(np.float32)
    tex = (tex.reshape(tex_size, tex_size, 3) * np.float32(tex_scale)).astype(np.float32)
    mats = [dict(kind="matte", kd=1.0, color=(1.0, 1.0, 1.0)),
            dict(kind="matte", kd=1.0, color=(0.5, 0.5, 0.5)),
            dict(kind="emissive", kd=1.0, color=(6.0, 6.0, 6.0))]
    shapes = [dict(kind="mesh", xyz=xyz, idx=idx, uv=uv, material=0, texture=0)]
    shapes += _quad((-6, -1.2, 6), (6, -1.2, 6), (6, -1.2, -6), (-6, -1.2, -6), 1, (0, 1, 0))
    # l x w points down (-y): the light faces the scene
    shapes.append(dict(kind="rect", p=(-1.0, 3.0, -1.0), l=(2.0, 0, 0), w=(0, 0, 2.0), material=2,
                       light=True))
    pos = (0.0, 0.6, 3.6)
    cam = dict(position=pos, basis=_look_at(pos, (0.0, 0.0, 0.0)), fov_deg=40.0, y_ratio=1.0)
    return dict(materials=mats, shapes=shapes, camera=cam, medium=None, textures=[tex])


def _look_at(position, target):
    """Row-major camera basis whose third column is the view direction."""
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    d /= np.linalg.norm(d)
    cx = np.cross(np.array([0.0, 1.0, 0.0]), d)
    cx /= np.linalg.norm(cx)
    cy = np.cross(d, cx)
    return np.stack([cx, cy, d], axis=1).tolist()


# ------------------------------------------------------------------ models
# Mat4x4 helpers with the reference's FP32 arithmetic (src/Kernel/matrix.cpp:
# Translation, Scaling, RotationX/Y/Z, operator *), row-major 4x4 float32.
def mat_translation(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[0, 3], m[1, 3], m[2, 3] = np.float32(x), np.float32(y), np.float32(z)
    return m


def mat_scaling(x, y, z):
    return np.diag(np.float32([x, y, z, 1.0]))


def _cs(radian):
    r = np.float32(radian)
    return np.float32(np.cos(np.float64(r))), np.float32(np.sin(np.float64(r)))


def mat_rotation_x(radian):
    c, s = _cs(radian)
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], np.float32)


def mat_rotation_y(radian):
    c, s = _cs(radian)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)


def mat_rotation_z(radian):
    c, s = _cs(radian)
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32)


def mat_mul(a, b):
    """Mat4x4::operator * : each entry summed left to right in FP32."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    out = np.empty((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            acc = a[i, 0] * b[0, j]
            for k in range(1, 4):
                acc = np.float32(acc + a[i, k] * b[k, j])
            out[i, j] = acc
    return out


# The model scene's transforms: RotationX(0) * Translation(0, -0.99, 0) *
# Scaling(s, s, s), s = 0.3 for the dragon and 0.65 for the teapot.
def model_transform(scale):
    return mat_mul(mat_mul(mat_rotation_x(0.0), mat_translation(0.0, -0.99, 0.0)),
                   mat_scaling(scale, scale, scale))


def model_scene(model: dict, width: int = 1200, height: int = 1000,
                absorption=(0.05, 0.25, 1.0)) -> dict:
    """The reference's model scene ("Build Model", src/Scene/Build Model.txt):
    a 6 x 5 x 6 room of Rect walls (white 0.7 floor, ceiling and back wall,
    green 0.67 right, red 0.67 left), a 3 x 3 Rect light of radiance 6.6666
    just under the ceiling, the model (load_obj's mesh fields) in a
    Dielectric(1.5, 1.00029, absorption) material, and a Pinhole at
    (0, 1.5, 20.2) turned by RotationY(pi), 20 degree field of view,
    y_ratio = height / width."""
    mats = [
        dict(kind="matte", kd=0.7, color=(1.0, 1.0, 1.0)),          # 0 floor, ceiling, back
        dict(kind="matte", kd=0.67, color=(0.01, 0.9, 0.01)),        # 1 right
        dict(kind="matte", kd=0.67, color=(0.9, 0.01, 0.01)),        # 2 left
        dict(kind="emissive", kd=6.6666, color=(1.0, 1.0, 1.0)),     # 3 light
        dict(kind="dielectric", eta_in=1.5, eta_out=1.00029, absorption=tuple(absorption)),  # 4
    ]
    shapes = [
        dict(kind="rect", p=(-3.0, -1.0, 3.0), l=(6.0, 0.0, 0.0), w=(0.0, 0.0, -6.0), material=0),
        dict(kind="rect", p=(3.0, -1.0, -3.0), l=(0.0, 0.0, 6.0), w=(0.0, 5.0, 0.0), material=1),
        dict(kind="rect", p=(-3.0, -1.0, 3.0), l=(0.0, 0.0, -6.0), w=(0.0, 5.0, 0.0), material=2),
        dict(kind="rect", p=(-3.0, 4.0, -3.0), l=(6.0, 0.0, 0.0), w=(0.0, 0.0, 6.0), material=0),
        dict(kind="rect", p=(-3.0, -1.0, -3.0), l=(6.0, 0.0, 0.0), w=(0.0, 5.0, 0.0), material=0),
        dict(kind="mesh", xyz=model["xyz"], idx=model["idx"], normals=model["normals"], material=4),
        dict(kind="rect", p=(-1.5, 3.999, -1.5), l=(3.0, 0.0, 0.0), w=(0.0, 0.0, 3.0), material=3,
             light=True),
    ]
    r = mat_rotation_y(np.pi)
    cam = dict(position=(0.0, 1.5, 20.2), basis=r[:3, :3].astype(np.float64).tolist(), fov_deg=20.0,
               y_ratio=height / width)
    return dict(materials=mats, shapes=shapes, camera=cam, medium=None)
