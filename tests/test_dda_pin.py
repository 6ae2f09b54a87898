"""Independent pin of the FP32 first-hit DDA (oracle or_render, which the
device marcher equals bit for bit): a float64 exact traversal
(oracle/dda_check.c) that derives each ray from the camera in float64 and
enumerates every LOD cell the ray passes through, with none of the FP32
walk's epsilons or per-brick re-initialisation (rayTrace.wgsl:457-563,
597-619, 647-847 are what the FP32 walk restates).

Bar, per frame: every pixel agrees with the exact first hit (label and
voxel index, t within the FP32 rounding bound) or is a grazing tie -- the
exact path stays less than that bound in some cell, and the FP32 answer is a
cell the other crossing order passes through. No pixel is unexplained, and
grazing ties stay below 0.1 % of the pixels. On exact ties inside a brick the
reference's tie rule (z, then y, then x) is required, with no allowance.
Planted mistakes in the FP32 oracle (swapped y/z tie order, a dropped entry
epsilon, a wrong x step sign) must produce unexplained pixels."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

from oracle import oracle as O

GRAZING_BAR = 1e-3


def orbit(G, yaw, pitch, radius, fov=math.pi / 3):
    """Orbit pose around the volume centre (orbitController.js:100-123) with
    the basis of camera.js:61-77, rounded to float32."""
    t = np.full(3, G / 2.0)
    p = t + radius * np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
    f = (t - p) / np.linalg.norm(t - p)
    wu = np.array([0.0, 1.0, 0.0]) if abs(f[1]) <= 0.999 else np.array([1.0, 0.0, 0.0])
    r = np.cross(f, wu)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    u /= np.linalg.norm(u)
    f32 = lambda v: [float(np.float32(x)) for x in v]
    return {"position": f32(p), "forward": f32(f), "right": f32(r), "up": f32(u),
            "tan_half_fov": float(np.float32(math.tan(fov / 2))), "ortho_half_height": 0.0}


def lod_rule(cam, B, b, W, H, fov=math.pi / 3):
    """calculateBrickLOD (camera.js:103-137) for every brick, Morton order."""
    L = b.bit_length() - 1
    out = np.zeros(B ** 3, np.uint8)
    for i in range(B ** 3):
        bx = sum(((i >> (3 * k)) & 1) << k for k in range(10))
        by = sum(((i >> (3 * k + 1)) & 1) << k for k in range(10))
        bz = sum(((i >> (3 * k + 2)) & 1) << k for k in range(10))
        c = ((bx + 0.5) * b, (by + 0.5) * b, (bz + 0.5) * b)
        out[i] = O.brick_lod(c, b, tuple(cam["position"]), fov, min(W, H), L)
    return out


class Scene:
    def __init__(self, data, b):
        self.ds = O.Dataset(data, b)
        self.b = b
        self.nb = self.ds.num_bricks
        self.B = round(self.nb ** (1 / 3))
        self.pyr, st = self.ds.decode_all_pyramid()
        assert st == 0
        self.empty = [int(np.all(self.ds.palette(i) == 0)) for i in range(self.nb)]
        self.p0 = [self.ds.palette(i)[0] for i in range(self.nb)]

    def labels(self, lods):
        return [self.pyr[i][O.pyramid_offset(int(lods[i])):O.pyramid_offset(int(lods[i]) + 1)]
                for i in range(self.nb)]

    def check(self, cam, lods, W, H):
        labels = self.labels(lods)
        lab, vox, tt, _ = O.render(self.B, self.b, self.empty, lods, labels, self.p0, cam, W, H)
        st, cls, _, _ = O.check_frame(self.B, self.b, self.empty, lods, labels, self.p0, cam, W, H, lab, vox, tt)
        return st


@pytest.fixture(scope="module")
def cfg1():
    with open(os.path.join(ROOT, "tests", "golden", "cfg1_b32.csdb"), "rb") as f:
        return Scene(f.read(), 32)


@pytest.fixture(scope="module")
def cfg2():
    import fcs

    return Scene(fcs.encode_generated(fcs.Generator(2), 64), 64)


def cfg1_cases(scene):
    G = scene.B * scene.b
    for k in range(8):
        cam = orbit(G, 0.3 + k * 0.77, -0.4 + 0.11 * k, (0.9 + 0.15 * k) * math.sqrt(3) * G)
        rng = np.random.default_rng(k)
        yield f"pose {k} LOD rule", cam, lod_rule(cam, scene.B, scene.b, 320, 240)
        yield f"pose {k} finest", cam, np.full(scene.nb, 5, np.uint8)
        yield f"pose {k} mixed", cam, rng.integers(0, 6, scene.nb).astype(np.uint8)


def test_perspective_frames_config1(cfg1):
    tot = {"pixels": 0, "grazing": 0, "unexplained": 0, "hits": 0}
    for name, cam, lods in cfg1_cases(cfg1):
        st = cfg1.check(cam, lods, 320, 240)
        assert st["unexplained"] == 0, (name, st)
        for k in tot:
            tot[k] += st[k]
    print(f"config 1: {tot['pixels']} pixels, {tot['hits']} hits, {tot['grazing']} grazing ties")
    assert tot["grazing"] <= GRAZING_BAR * tot["pixels"], tot


def test_perspective_frames_config2(cfg2):
    """The 512^3 demo volume (64^3 bricks) from the benchmark's orbit poses
    (radius 1.5x the diagonal, yaw pi/4 + k pi/4, pitch 0.15 pi), at the LOD
    rule and at the finest LOD, plus two poses from inside the volume."""
    G = cfg2.B * cfg2.b
    W, H = 480, 270
    tot = {"pixels": 0, "grazing": 0, "unexplained": 0, "hits": 0}
    cams = [orbit(G, math.pi / 4 + k * math.pi / 4, 0.15 * math.pi, 1.5 * math.sqrt(3) * G) for k in range(0, 8, 2)]
    cams += [orbit(G, 1.1, -0.2, 0.3 * G), orbit(G, 4.0, 0.5, 0.45 * G)]
    for i, cam in enumerate(cams):
        for lods in (lod_rule(cam, cfg2.B, cfg2.b, W, H), np.full(cfg2.nb, 6, np.uint8)):
            st = cfg2.check(cam, lods, W, H)
            assert st["unexplained"] == 0, (i, st)
            for k in tot:
                tot[k] += st[k]
    print(f"config 2: {tot['pixels']} pixels, {tot['hits']} hits, {tot['grazing']} grazing ties")
    assert tot["hits"] > 0.2 * tot["pixels"]
    assert tot["grazing"] <= GRAZING_BAR * tot["pixels"], tot


# ------------------------------------------------------------ exact ties --
S = float(np.float32(1.0 / math.sqrt(2.0)))


def tie_scene(cell_label):
    """One 16^3 brick at its finest LOD, labels from cell_label(x, y, z)."""
    b = 16
    labels = np.zeros(b ** 3, np.uint64)
    for z in range(b):
        for y in range(b):
            for x in range(b):
                m = 0
                for k in range(4):
                    m |= ((x >> k) & 1) << (3 * k) | ((y >> k) & 1) << (3 * k + 1) | ((z >> k) & 1) << (3 * k + 2)
                labels[m] = cell_label(x, y, z)
    return labels


def one_ray(origin, direction, labels):
    """A 1x1 orthographic frame: its single ray starts at `origin` exactly and
    runs along `direction` (float32 values, so the float64 check sees the
    same ray)."""
    cam = {"position": [float(v) for v in origin], "forward": [float(v) for v in direction],
           "right": [1.0, 0.0, 0.0], "up": [0.0, 1.0, 0.0], "tan_half_fov": 0.0, "ortho_half_height": 1.0}
    lab, vox, tt, _ = O.render(1, 16, [0], [4], [labels], [0], cam, 1, 1)
    st, cls, el, ev = O.check_frame(1, 16, [0], [4], [labels], [0], cam, 1, 1, lab, vox, tt)
    return st, int(lab[0]), int(vox[0])


# (axis pair, direction, start, step-order cell offsets): the ray crosses the
# two axes' planes at exactly the same parameters; `rule` is the cell the
# reference samples between two tied crossings, `other` the one it must not.
TIE_CASES = [
    ("y/z", (0.0, S, S), (5.5, 0.25, 2.25), 2, 1),
    ("y/z negative", (0.0, -S, -S), (5.5, 12.75, 14.75), 2, 1),
    ("x/y", (S, S, 0.0), (1.25, 3.25, 7.5), 1, 0),
    ("x/z", (S, 0.0, S), (2.25, 7.5, 0.25), 2, 0),
]


@pytest.mark.parametrize("name,d,o,first,second", TIE_CASES, ids=[c[0] for c in TIE_CASES])
def test_exact_ties_follow_the_rule(name, d, o, first, second):
    """Labels only on the cells a tied step reaches through the axis the rule
    steps first: the FP32 walk must hit the first one; labels only on the
    cells the other order would visit: it must hit nothing."""
    c0 = [int(math.floor(v)) for v in o]
    sgn = [1 if v > 0 else (-1 if v < 0 else 0) for v in d]

    def cells(axis_first):
        # cells entered between two tied crossings when `axis_first` steps first
        out = set()
        c = list(c0)
        for _ in range(16):
            c[axis_first] += sgn[axis_first]
            out.add(tuple(c))
            other = second if axis_first == first else first
            c[other] += sgn[other]
        return out

    rule_cells, other_cells = cells(first), cells(second)
    lab_rule = tie_scene(lambda x, y, z: 7 if (x, y, z) in rule_cells else 0)
    st, lab, vox = one_ray(o, d, lab_rule)
    assert st["agree"] == 1 and lab == 7, (name, st)
    exp = [c0[k] + (sgn[k] if k == first else 0) for k in range(3)]
    assert vox == exp[0] + 16 * (exp[1] + 16 * exp[2])
    lab_other = tie_scene(lambda x, y, z: 9 if (x, y, z) in other_cells else 0)
    st, lab, _ = one_ray(o, d, lab_other)
    assert st["agree"] == 1 and lab == 0, (name, st)


# ---------------------------------------------------- planted mistakes --
DDA_MUTANTS = {7: "y/z tie order swapped", 8: "entry epsilon dropped", 9: "x step sign wrong"}


def _with_mutant(mutant, tmp_path, fn):
    so = tmp_path / f"liboracle_m{mutant}.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    f"-DOR_MUTANT={mutant}", "-o", str(so), os.path.join(ROOT, "oracle", "oracle.c"), "-lm",
                    "-lpthread"], check=True)
    saved, real = O._lib, O.ORACLE_SO
    try:
        O._lib = None
        O.ORACLE_SO = str(so)
        return fn()
    finally:
        O.ORACLE_SO = real
        O._lib = saved


@pytest.mark.parametrize("mutant", sorted(DDA_MUTANTS))
def test_planted_dda_mistakes_are_caught(mutant, cfg1, tmp_path):
    def run():
        unexplained = 0
        for _, cam, lods in cfg1_cases(cfg1):
            unexplained += cfg1.check(cam, lods, 160, 120)["unexplained"]
        for case in TIE_CASES:
            try:
                test_exact_ties_follow_the_rule(*case)
            except AssertionError:
                unexplained += 1
        return unexplained

    assert _with_mutant(mutant, tmp_path, run) > 0, DDA_MUTANTS[mutant]
