"""The oracle's progressive display step (o_display_frame, restating
RenderKernel.cu:202-217 and Renderer.cpp:86-93) against closed forms: the 8-bit value
floor(255 * c^(1/2.2)) computed in float64, the clamp ends, NaN, the alpha byte, the
running sum's byte-1 starting value and the division by the frame count. Every check
is run again on oracle builds with a planted mistake (rounding instead of truncation,
a sum starting from 0, exponent 2.2 instead of 1/2.2), and at least one must fail."""
import numpy as np
import pytest

from oracle import oracle

FRESH = np.array([0x01010101], np.uint32).view(np.float32)[0]   # cudaMemset(buffer, 1, ...)


def _channels(rgba):
    return np.stack([(rgba >> (8 * k)) & 0xFF for k in range(4)], -1)


def _gamma8(mean):
    """floor(255 * clamp(mean)^(1/2.2)) in float64, and whether it is far enough from an
    integer boundary that float32 rounding cannot move it."""
    c = np.clip(np.asarray(mean, np.float64), 0.0, 1.0)
    x = 255.0 * c ** (1.0 / 2.2)
    return np.floor(x).astype(np.int64), np.abs(x - np.round(x)) > 1e-3


def _failures(lib_):
    fails = []
    rng = np.random.default_rng(7)
    # one frame of values spread over [0, 1]
    c = rng.random((4096, 3), dtype=np.float32)
    s = np.zeros_like(c)
    rgba = oracle.display_frame(c, s, 1, lib_=lib_)
    ch = _channels(rgba)
    want, ok = _gamma8(c.astype(np.float64) + float(FRESH))
    if not np.array_equal(ch[..., :3][ok], want[ok]):
        fails.append("gamma8 of one frame")
    if not np.all(ch[..., 3] == 1):
        fails.append("alpha byte")
    # clamp ends and non-finite values (fmaxf(0, fminf(x, 1)): NaN -> 1)
    ends = np.array([[-1.0, 0.0, 1.0], [2.0, np.inf, -np.inf], [np.nan, 1e30, -0.0]], np.float32)
    s = np.zeros_like(ends)
    ch = _channels(oracle.display_frame(ends, s, 1, lib_=lib_))[..., :3]
    if not np.array_equal(ch, np.array([[0, 0, 255], [255, 255, 0], [255, 255, 0]])):
        fails.append(f"clamp ends {ch.tolist()}")
    # the running sum starts at the byte-1 pattern on frame 1, whatever the buffer held
    z = np.zeros((16, 3), np.float32)
    s = np.full_like(z, 123.0)
    ch = _channels(oracle.display_frame(z, s, 1, lib_=lib_))[..., :3]
    if not (np.all(s.view(np.uint32) == 0x01010101) and np.all(ch == 0)):
        fails.append("byte-1 starting sum")
    # three frames: float32 running sum in order, mean = sum / 3
    frames = [rng.random((2048, 3), dtype=np.float32) * 1.5 for _ in range(3)]
    s = np.zeros_like(frames[0])
    want_sum = np.full_like(frames[0], FRESH)
    for k, f in enumerate(frames, 1):
        rgba = oracle.display_frame(f, s, k, lib_=lib_)
        want_sum = (want_sum + f).astype(np.float32)
    if not np.array_equal(s.view(np.uint32), want_sum.view(np.uint32)):
        fails.append("running sum")
    want, ok = _gamma8(want_sum.astype(np.float64) / 3.0)
    if not np.array_equal(_channels(rgba)[..., :3][ok], want[ok]):
        fails.append("gamma8 of the three-frame mean")
    return fails


def test_display_closed_forms():
    assert _failures(oracle.lib) == []


@pytest.mark.parametrize("plant", [16, 17, 18])
def test_display_checks_catch_planted_mistakes(plant):
    assert _failures(oracle.load_planted(plant)), f"planted mistake {plant} passed every display check"
