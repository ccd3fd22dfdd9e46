"""Pins the oracle's ImageBasedLight (src/Light/imagebasedlight.cpp:35-64)
on the CPU by closed forms: a matte surface under a sky of constant
radiance E gets rho * E from every light sample (cosine sampling cancels
the cosine and the 1/pi of the pdf), and the sky seen directly is E."""
import numpy as np
import pytest

import render_cases as rc


@pytest.mark.parametrize("render", ["pt", "whitted"])
def test_constant_sky_lights_a_matte_floor_exactly(oracle, render):
    E = np.array([1.0, 2.0, 3.0])
    sc = rc.ibl_scene(8, 8, constant=E)
    sc["camera"] = dict(position=(0.0, 1.0, 0.3), basis=rc.look_at_basis((0, 1, 0.3), (0, 0, 0)),
                        fov_deg=0.5, y_ratio=1.0)
    fn = oracle.render_pt if render == "pt" else oracle.render_whitted
    mean, m2 = fn(sc, 2, 2, 64, 3, seed=2, nthreads=1)
    np.testing.assert_allclose(mean, np.broadcast_to(0.6 * E, mean.shape), rtol=1e-6)
    np.testing.assert_allclose(m2, mean * mean, rtol=1e-6)      # no variance


def test_sky_seen_from_inside_is_its_texture(oracle):
    E = np.array([0.5, 1.5, 2.5])
    sc = rc.ibl_scene(8, 8, constant=E)
    sc["camera"] = dict(position=(0.0, 1.0, 0.0), basis=rc.look_at_basis((0, 1, 0), (0.3, 3.0, 0.2)),
                        fov_deg=1.0, y_ratio=1.0)
    mean, _ = oracle.render_pt(sc, 2, 2, 16, 3, seed=1, nthreads=1)
    np.testing.assert_allclose(mean, np.broadcast_to(E, mean.shape), rtol=1e-9)


def test_image_based_light_is_path_tracer_and_raytracer_only(oracle):
    sc = rc.ibl_scene(8, 8)
    with pytest.raises(RuntimeError):
        oracle.render_pt(dict(sc, medium=dict(bound="box", lo=(-1, 0, -1), hi=(1, 1, 1),
                                              sigma_s=(0.1,) * 3, sigma_a=(0.1,) * 3)),
                         8, 8, 4, 3, seed=1, nthreads=1)
