"""Pins of the oracle's output formatting and vertex colours (CPU only).

Colours (ColorMapper::generate_colors, ColorMapper.cpp:15-119): the HSV map
against the standard library's textbook conversion (colorsys) and hand
values; the colour noise is the pinned terrain FBM (a terrain sampler with
amp 1 and y0 = y returns the FBM exactly).  Formatting (GLChunk::format_data,
DMCChunk::copy_verts_and_inds): chunks without a mesh are skipped, indices
are offset by the chunk's first formatted vertex, world positions are
p * scale + overlap_pos.
"""
import colorsys
import math

import numpy as np
import pytest


def test_hsv_hand_values(oracle):
    f = np.float32
    p = f(1.0) * (f(1.0) - f(0.72))
    assert np.array_equal(oracle.hsv_to_rgb(0.0, 0.72, 1.0), np.array([1.0, p, p], np.float32))
    assert np.array_equal(oracle.hsv_to_rgb(120.0, 0.72, 1.0), np.array([p, 1.0, p], np.float32))
    assert np.array_equal(oracle.hsv_to_rgb(240.0, 0.72, 1.0), np.array([p, p, 1.0], np.float32))
    assert np.array_equal(oracle.hsv_to_rgb(360.0, 0.72, 1.0), np.array([1.0, p, p], np.float32))
    # negative hues are folded by fabs before the modulo (ColorMapper.cpp:78)
    assert np.array_equal(oracle.hsv_to_rgb(-120.0, 0.72, 1.0), oracle.hsv_to_rgb(120.0, 0.72, 1.0))


def test_hsv_matches_colorsys(oracle):
    rng = np.random.default_rng(5)
    for h in np.concatenate([rng.uniform(-900.0, 900.0, 2000), np.arange(0.0, 720.0, 7.5)]):
        hh = math.fmod(abs(float(np.float32(h))), 360.0) / 360.0
        want = colorsys.hsv_to_rgb(hh, 0.72, 1.0)
        got = oracle.hsv_to_rgb(float(h), 0.72, 1.0)
        assert np.abs(got - np.array(want)).max() < 2e-6, (h, got, want)


def test_color_noise_is_the_terrain_fbm(oracle):
    rng = np.random.default_rng(9)
    for x, y, z in rng.uniform(-300.0, 300.0, (500, 3)).astype(np.float32):
        spec = oracle.SamplerSpec(kind=oracle.SAMPLER_TERRAIN, amp=1.0, y0=float(y), freq=0.01, lacunarity=2.0,
                                  gain=0.5, octaves=4, seed=1337)
        want = np.float32(oracle.lib().or_terrain_value_f32(spec.ref, float(x), float(y), float(z)))
        assert np.float32(oracle.color_noise(float(x), float(y), float(z))) == want


def test_vertex_color_composition(oracle):
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, 64.0, (200, 3)).astype(np.float32)
    col = oracle.vertex_colors(pos)
    for p, c in zip(pos, col):
        n = np.float32(oracle.color_noise(*map(float, p))) * np.float32(4.0)
        h = ((n + np.float32(1.0)) * np.float32(0.5)) * np.float32(360.0)
        assert np.array_equal(c, oracle.hsv_to_rgb(float(h), 0.72, 1.0))
    assert np.all((col >= 0.28 - 1e-6) & (col <= 1.0))


def _chunk(rng, V, I, cm, delta, opos):
    return dict(V=V, I=I, contains_mesh=cm, delta=delta, opos=opos,
                pos=rng.uniform(0, 32, (V, 3)).astype(np.float32),
                nrm=rng.standard_normal((V, 3)).astype(np.float32),
                idx=rng.integers(0, max(V, 1), I).astype(np.uint32))


@pytest.mark.parametrize("space", [0, 1])
@pytest.mark.parametrize("glob", [False, True])
def test_format_batch_layout(oracle, space, glob):
    rng = np.random.default_rng(1)
    chunks = [_chunk(rng, 10, 24, True, 2.0, (64.0, 0.0, -8.0)),
              _chunk(rng, 7, 12, False, 1.0, (0.0, 0.0, 0.0)),    # no mesh: skipped
              _chunk(rng, 0, 0, True, 1.0, (0.0, 0.0, 0.0)),
              _chunk(rng, 5, 9, True, 0.5, (1.5, 2.5, 3.5))]
    pos, col, nrm, idx = oracle.format_batch(chunks, space, 1, glob)
    kept = [c for c in chunks if c["contains_mesh"]]
    assert len(pos) == sum(c["V"] for c in kept) and len(idx) == sum(c["I"] for c in kept)
    v0 = i0 = 0
    for c in kept:
        p = c["pos"]
        if space == 1:
            p = (p * np.float32(c["delta"])) + np.array(c["opos"], np.float32)
        assert np.array_equal(pos[v0:v0 + c["V"]], p)
        assert np.array_equal(nrm[v0:v0 + c["V"]], c["nrm"])
        assert np.array_equal(idx[i0:i0 + c["I"]], c["idx"] + (np.uint32(v0) if glob else np.uint32(0)))
        v0 += c["V"]
        i0 += c["I"]
    assert np.all(col == 1.0)
