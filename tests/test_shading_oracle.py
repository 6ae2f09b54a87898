"""Shaded (RGBA) frames on the CPU: the transfer function colours against an
independent HSL computation, and the compositing oracle (or_render_rgba)
against closed forms on volumes whose composite is known analytically."""
import colorsys
import math
import os
import subprocess

import numpy as np
import pytest

import fcs
from conftest import ROOT
from fcs._lib import lib
from oracle import oracle as O


def js_round(x):
    f = math.floor(x)
    return f + 1 if x - f >= 0.5 else f


def tf_colour_closed_form(i):
    """webgpuSetup.js:410-429 with the textbook HSL->RGB (colorsys)."""
    hue = (i * 137.5) % 360.0
    s = 0.55 + (i % 3) * 0.08
    l = 0.35 + (i % 4) * 0.05
    a = 0.25 + (i % 5) * 0.08
    r, g, b = colorsys.hls_to_rgb(hue / 360.0, l, s)
    ch = [min(math.floor(js_round(c * 255) * 0.92), 255) for c in (r, g, b)]
    return ch[0] | (ch[1] << 8) | (ch[2] << 16) | (math.floor(a * 255) << 24)


def test_transfer_function_colours():
    for i in list(range(400)) + [1000, 4095, 65536, 1 << 20, 2_000_003]:
        assert lib.fcs_tf_default_rgba(i) == tf_colour_closed_form(i), i
    # opacities cycle through 0.25 .. 0.57
    assert [lib.fcs_tf_default_rgba(i) >> 24 for i in range(5)] == [63, 84, 104, 124, 145]


def _scene(labels_zyx, b, lod):
    """Brick arrays of a small volume at one LOD for every brick."""
    D = labels_zyx.shape[0]
    B = D // b
    nb = B ** 3
    empty = np.zeros(nb, np.uint8)
    lods = np.full(nb, lod, np.uint8)
    p0 = np.zeros(nb, np.uint64)
    bricks = []
    n = 1 << lod
    s = b // n
    idx = np.arange(n ** 3)
    cx = sum(((idx >> (3 * k)) & 1) << k for k in range(10))
    cy = sum(((idx >> (3 * k + 1)) & 1) << k for k in range(10))
    cz = sum(((idx >> (3 * k + 2)) & 1) << k for k in range(10))
    for i in range(nb):
        bx = sum(((i >> (3 * k)) & 1) << k for k in range(10))
        by = sum(((i >> (3 * k + 1)) & 1) << k for k in range(10))
        bz = sum(((i >> (3 * k + 2)) & 1) << k for k in range(10))
        # a coarse cell takes the label of its min-corner voxel (a uniform region keeps its label)
        lab = labels_zyx[bz * b + cz * s, by * b + cy * s, bx * b + cx * s].astype(np.uint64)
        bricks.append(lab)
        empty[i] = 1 if not np.any(labels_zyx[bz * b:(bz + 1) * b, by * b:(by + 1) * b, bx * b:(bx + 1) * b]) else 0
        p0[i] = lab[0]
    return B, nb, empty, lods, p0, bricks


def _ortho_z(G):
    return {"position": [G / 2, G / 2, -1.0], "forward": [0, 0, 1], "right": [1, 0, 0], "up": [0, -1, 0],
            "tan_half_fov": 0.0, "ortho_half_height": G / 2}


def _unpack(px):
    return [(int(px) >> (8 * k)) & 0xFF for k in range(4)]


def _over(samples):
    """Front-to-back 'over' of (r, g, b, a) samples until alpha > 0.95."""
    C = [0.0, 0.0, 0.0]
    A = 0.0
    for r, g, b, a in samples:
        src = a * (1 - A)
        C = [C[0] + r * src, C[1] + g * src, C[2] + b * src]
        A += src
        if A > 0.95:
            break
    return C, A


@pytest.mark.parametrize("lod", [4, 2])
def test_uniform_volume_composite(lod):
    """One label everywhere, seen along +z: every sample has alpha a (below the
    shading threshold) and opacity correction 1-(1-a)^stride, so the pixel is
    the label colour times A = 1-(1-a)^(stride*n), n the first count above 0.95."""
    G, b, L = 64, 16, 7777
    vol = np.full((G, G, G), L, np.uint64)
    B, nb, empty, lods, p0, bricks = _scene(vol, b, lod)
    rgba8 = 0x3f_40_80_c0  # alpha 63/255 < 0.35: no shading
    out = O.render_rgba(B, b, empty, lods, bricks, p0, _ortho_z(G), 16, 16, [L], [rgba8], [0.4, 0.6, -0.9])
    a = 63 / 255
    stride = b >> lod
    n = math.ceil(math.log(0.05) / (stride * math.log(1 - a)))
    A = 1 - (1 - a) ** (stride * n)
    want = [0xC0 / 255 * A, 0x80 / 255 * A, 0x40 / 255 * A, A]
    for px in out:
        got = _unpack(px)
        for k in range(4):
            assert abs(got[k] - want[k] * 255) <= 1.0, (lod, got, [w * 255 for w in want])


def test_interface_normal_and_diffuse():
    """Label B (alpha 15/255) for z < 32, label A (alpha 97/255: shaded, no
    shadow ray) for z >= 32, seen along +z. At the interface the gradient is
    along z only, so the normal is (0, 0, -1) and the shading factor is
    ambient + diffuse_strength * (n . l/|l|); inside A the gradient vanishes
    (factor = ambient)."""
    G, b = 64, 16
    LA, LB = 5, 9
    vol = np.full((G, G, G), LB, np.uint64)
    vol[32:] = LA
    B, nb, empty, lods, p0, bricks = _scene(vol, b, 4)
    ca, cb = 0x61_20_60_a0, 0x0f_30_30_30
    light = np.array([math.sin(math.radians(25)), 0.6, -math.cos(math.radians(25))], np.float32)
    amb, dif = 0.2, 0.5
    out = O.render_rgba(B, b, empty, lods, bricks, p0, _ortho_z(G), 8, 8, [LA, LB], [ca, cb], light, amb, dif)
    rgb = lambda c: [(c >> (8 * k) & 0xFF) / 255 for k in range(3)]
    aA, aB = 97 / 255, 15 / 255
    ndotl = float(-light[2] / np.linalg.norm(light.astype(np.float64)))
    f_if = min(max(amb + ndotl * dif, 0.3), 1.0)
    f_in = max(amb, 0.1)
    samples = [(*rgb(cb), aB)] * 32 + [(*[c * f_if for c in rgb(ca)], aA)] + [(*[c * f_in for c in rgb(ca)], aA)] * 32
    C, A = _over(samples)
    want = [C[0], C[1], C[2], A]
    for px in out:
        got = _unpack(px)
        for k in range(4):
            assert abs(got[k] - want[k] * 255) <= 1.0, (got, [w * 255 for w in want])


def test_hidden_label_is_transparent():
    G, b, L = 32, 16, 12
    vol = np.full((G, G, G), L, np.uint64)
    B, nb, empty, lods, p0, bricks = _scene(vol, b, 4)
    out = O.render_rgba(B, b, empty, lods, bricks, p0, _ortho_z(G), 8, 8, [L], [0], [0.4, 0.6, -0.9])
    assert np.all(out == 0)


def test_planted_opacity_correction_mistake_is_caught(tmp_path):
    so = tmp_path / "liboracle_m6.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", "-DOR_MUTANT=6", "-o", str(so),
                    os.path.join(ROOT, "oracle", "oracle.c"), "-lm", "-lpthread"], check=True)
    saved, real = O._lib, O.ORACLE_SO
    try:
        O._lib = None
        O.ORACLE_SO = str(so)
        with pytest.raises(AssertionError):
            test_uniform_volume_composite(2)
    finally:
        O._lib, O.ORACLE_SO = saved, real
