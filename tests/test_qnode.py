"""The traversal's 64-byte quantized BVH nodes (cpt_debug_quantize_node, host code,
no GPU): every quantized child box contains its exact float box (checked in exact
rational arithmetic), absent children get empty boxes, the scale is the smallest
power of two that spans the node, and the boxes grow by at most two scale steps."""
import ctypes as C
from fractions import Fraction

import numpy as np

import cpt

EMPTY = 0x7FFFFFFF
EMIN, EMAX = -100, 60


def _quantize(lib, lo, hi, child):
    """lo, hi: 3 x 4 float32 (axis, child); child: 4 int32."""
    n4 = np.zeros(32, np.uint32)
    planes = np.zeros((6, 4), np.float32)
    for a in range(3):
        planes[2 * a], planes[2 * a + 1] = lo[a], hi[a]
    n4[:24] = planes.reshape(-1).view(np.uint32)
    n4[24:28] = np.asarray(child, np.int32).view(np.uint32)
    out = np.zeros(16, np.uint32)
    assert lib.cpt_debug_quantize_node(n4.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)) == 0
    org = out[:3].view(np.float32)
    exps = [int((out[3] >> (8 * a)) & 255) - 127 for a in range(3)]
    kids = out[4:8].view(np.int32)
    qlo = [[int((out[8 + 2 * a] >> (8 * c)) & 255) for c in range(4)] for a in range(3)]
    qhi = [[int((out[9 + 2 * a] >> (8 * c)) & 255) for c in range(4)] for a in range(3)]
    return org, exps, kids, qlo, qhi


def _random_node(rng, mode):
    off = 10.0 ** rng.uniform(-8, 6) * rng.choice([-1.0, 1.0])
    size = 10.0 ** rng.uniform(-6, 4)
    lo = np.zeros((3, 4), np.float32)
    hi = np.zeros((3, 4), np.float32)
    child = np.arange(4, dtype=np.int32)
    if mode == 3:
        child[2] = EMPTY
    for a in range(3):
        for c in range(4):
            low = off + size * (rng.random() - 0.5) * 4
            w = 0.0 if mode == 1 else size * 1e-7 * rng.random() if mode == 4 else size * rng.random()
            lo[a, c] = np.float32(low)
            hi[a, c] = max(np.float32(low + w), lo[a, c])
        if mode == 2:  # equal boxes
            lo[a, 1:], hi[a, 1:] = lo[a, 0], hi[a, 0]
    if mode == 3:
        lo[:, 2], hi[:, 2] = np.inf, -np.inf
    return lo, hi, child


def test_quantized_boxes_contain_the_exact_boxes():
    lib = cpt.load_library()
    rng = np.random.default_rng(5)
    worst = 0.0
    for it in range(3000):
        lo, hi, child = _random_node(rng, it % 5)
        org, exps, kids, qlo, qhi = _quantize(lib, lo, hi, child)
        assert list(kids) == list(child)
        for a in range(3):
            present = [c for c in range(4) if child[c] != EMPTY]
            o = Fraction(float(org[a]))
            assert org[a] == min(lo[a, c] for c in present)  # the children's low corner, exactly
            k = exps[a]
            assert EMIN <= k <= EMAX
            s = Fraction(2) ** k
            ext = max(Fraction(float(hi[a, c])) for c in present) - o
            assert 255 * s >= ext  # the scale spans the node
            if k > EMIN:  # and is the smallest power of two that does (with its 1e-6 margin)
                assert 255 * s / 2 < ext * Fraction(1000001, 1000000)
            for c in range(4):
                if child[c] == EMPTY:
                    assert (qlo[a][c], qhi[a][c]) == (255, 0)
                    continue
                plo, phi = o + qlo[a][c] * s, o + qhi[a][c] * s
                assert plo <= Fraction(float(lo[a, c])) and phi >= Fraction(float(hi[a, c])), (it, a, c)
                # outward by less than two steps on each side
                assert Fraction(float(lo[a, c])) - plo < 2 * s and phi - Fraction(float(hi[a, c])) < 2 * s
                if ext > 0:
                    worst = max(worst, float((Fraction(float(lo[a, c])) - plo + phi - Fraction(float(hi[a, c]))) / ext))
    assert worst < 0.02  # measured about 0.016 of the node's extent


def test_quantization_rounds_outward_at_exact_steps():
    """A child plane exactly on a step keeps it; one float above (below) a step is
    rounded down (up) to the step outside, never to the nearest one."""
    lib = cpt.load_library()
    lo = np.zeros((3, 4), np.float32)
    hi = np.zeros((3, 4), np.float32)
    hi[:, 0] = 255.0  # node extent 255: scale 2 with the margin (1 x 255 is not above 255)
    lo[:, 1], hi[:, 1] = 4.0, 6.0  # on steps of 2
    lo[:, 2], hi[:, 2] = np.nextafter(np.float32(4.0), np.float32(5.0)), np.nextafter(np.float32(6.0), np.float32(5.0))
    lo[:, 3], hi[:, 3] = np.float32(3.0), np.float32(7.0)  # between steps
    org, exps, _, qlo, qhi = _quantize(lib, lo, hi, [0, 1, 2, 3])
    assert list(org) == [0.0, 0.0, 0.0] and exps == [1, 1, 1]
    for a in range(3):
        assert (qlo[a][1], qhi[a][1]) == (2, 3)
        assert (qlo[a][2], qhi[a][2]) == (2, 3)  # 4 + ulp rounds down to 4, 6 - ulp up to 6
        assert (qlo[a][3], qhi[a][3]) == (1, 4)  # 3 -> 2, 7 -> 8
