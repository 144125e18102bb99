"""A Mandelbulber scene with a closed-form field: the Kaleidoscopic IFS
(Shaders.metal:897-949) with no folds, offset or rotation maps z -> k z, so
with N iterations and bailout B > 2 its distance inside the non-escaping
ball |z| <= B / k^N is exactly |p| - R, R = 2 ws / k^N (128 world units
here), and a lower bound outside.  Camera at (0, 0, -1000) looking along +z.
"""
from fractalpathtracer_amd import ffi, scenes

WS, K, N = 1024.0, 2.0, 4
R = 2 * WS / K ** N  # 128 world units


def sphere_cfg(mode=0, thr=2.0, factor=0.002, bailout=100.0, detail=2.0):
    c = scenes.default_config()
    c.sdf_id = ffi.SDF_MANDELBULBER
    sv = c.set_values
    for i, x in enumerate([WS, N, bailout, K, 0, 0, 0, 0, 0, 0, 0, 0]):
        sv[i] = x
    v = c.vset_values
    for i in range(107, 131):
        v[i] = 0.0
    v[110], v[111], v[112], v[113], v[114] = 3072.0, 1.0, 10000.0, 1.0, 1.0
    v[115], v[116], v[117], v[118], v[119], v[129] = mode, factor, thr, 0.0, 1.0e9, detail
    c.camera_position[:] = [0.0, 0.0, -1000.0]
    c.camera_yaw_pitch[:] = [0.0, 0.0]
    c.camera_roll = 0.0
    c.camera_fov = 60.0
    c.width, c.height = 64, 32
    c.render[1], c.render[4], c.render[6], c.render[7] = 10000.0, 51200.0, 0.0, 0.0
    return c

LIMIT = 1 - 0.001 * 2.0  # the refinement band's lower edge at detail level 2
