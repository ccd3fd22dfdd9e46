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
