"""Scenes whose path-tracer estimate has a closed form or exercises the
estimator's rarely taken branches (renderPath, Shaders.metal:5926-6026).

cavity_cfg: the camera sits at the centre of a spherical cavity (a box minus a
sphere of radius 1), so every march of every bounce ends on the cavity wall
and no path reaches the environment.  With a constant diffuse material
(roughness 1, specular 0, translucency 0, albedo a, emission e) and no sun,
each bounce adds a*e to the light sum and multiplies the throughput by a, and
the estimator returns light * throughput (SM:6024): after N bounces the
sample is N a e a^N per channel, clamped at 8 (SM:6026).

glass_cfg: a translucent, emissive sphere resting on a translucent box under
the sun and a constant sky, for parity of the refraction, Fresnel, emission
and translucency-LOD branches.
"""
from fractalpathtracer_amd import ffi, scenes


def cavity_cfg(bounces=3, albedo=(0.9, 0.6, 0.3), emission=1.5, width=48, height=40, spp=2,
               bounce_cap=0, rr=False, rr_start=2.0, rr_min=0.05):
    c = scenes.default_config()
    c.width, c.height, c.samples = width, height, spp
    c.camera_position[:] = [0.0, 0.0, 0.0]
    c.camera_yaw_pitch[:] = [0.3, -0.2]
    c.camera_fov, c.camera_dof, c.camera_roll = 70.0, 0.0, 0.0
    c.render[:] = [float(bounces), 256.0, 0.0005, 1e-4, 100.0, 0.0, 0.0, 0.0]
    c.sun[0] = 0.0
    c.world[4] = 0.0
    c.world[6] = 0.0
    scenes.set_program(c, [(ffi.OP_BOX, 0, [2.0, 2.0, 2.0, 0.0]), (ffi.OP_SPHERE, 2, [1.0, 0.0, 0.0, 0.0])])
    c.material_mode = 0
    c.program_material[:] = [*albedo, 1.0, 0.0, 0.0, 1.5, emission]
    c.sdf_bounce_cap = bounce_cap
    c.sdf_russian_roulette = 1 if rr else 0
    c.sdf_rr_start, c.sdf_rr_min_prob = rr_start, rr_min
    return c


def glass_cfg(width=64, height=48, spp=4, bounces=5, translucency=0.7, emission=0.25, min_distance=1e-3):
    c = scenes.default_config()
    c.width, c.height, c.samples = width, height, spp
    scenes.look_at(c, (0.6, 0.9, -3.2), (0.0, -0.2, 0.0))
    c.camera_fov, c.camera_dof, c.camera_roll = 55.0, 0.0, 0.0
    c.render[:] = [float(bounces), 256.0, 0.0005, min_distance, 50.0, 0.0, 0.0, 0.0]
    c.sun[:] = [1.0, scenes.SUN_ROTATION_DEG, scenes.SUN_ELEVATION_DEG, 1.5, 0.0]
    c.sun_color[:] = [1.0, 0.95, 0.9]
    c.world[:] = [3.0, 1.0, 0.0, 0.0, 0.6, 1.0, 0.0]
    c.world_one_color[:] = [0.5, 0.6, 0.8]
    scenes.set_program(c, [(ffi.OP_SPHERE, 0, [0.7, 0.0, 0.0, 0.0]),
                           (ffi.OP_TRANSLATE, 0, [0.0, -1.0, 0.0, 0.0]),
                           (ffi.OP_BOX, 0, [2.5, 0.25, 2.5, 0.0])])
    c.material_mode = 0
    c.program_material[:] = [0.85, 0.55, 0.35, 0.3, 0.5, translucency, 1.45, emission]
    c.sdf_russian_roulette = 1
    c.sdf_rr_start, c.sdf_rr_min_prob = 3.0, 0.2
    return c
