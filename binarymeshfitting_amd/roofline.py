"""Algorithmic work of the extraction kernels, for roofline fractions.

FLOP constants are read from csrc/samplers.h (counted from the kernel
source there).  MI355X peaks are the published figures.
"""
import re
from pathlib import Path

_SAMPLERS = Path(__file__).resolve().parent / "csrc" / "samplers.h"
_TEXT = _SAMPLERS.read_text()


def _const(name):
    return int(re.search(rf"constexpr int {name} = (\d+);", _TEXT).group(1))


SIMPLEX_FLOPS = _const("kSimplexFlops")
OCTAVE_FLOPS = _const("kOctaveFlops")
TERRAIN_FLOPS = _const("kTerrainFlops")
HASH_INT_OPS = _const("kHashIntOps")
CSG_SPHERE_FLOPS = _const("kCsgSphereFlops")
CSG_BOX_FLOPS = _const("kCsgBoxFlops")

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector, packed (78.6 unpacked)
PEAK_FP32_UNPACKED_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0      # HBM3E

# Compulsory HBM bytes written by the sampling kernel per sample: one label
# bit.  Densities are stored only where a vertex needs them (both ends of a
# sign-changing edge), a data-dependent fraction that is not counted here.
LABEL_BYTES = 1.0 / 8.0


def terrain_flops_per_sample(octaves):
    """FP32 operations of one full-octave density evaluation (the reference
    evaluates every octave at every sample)."""
    return octaves * (SIMPLEX_FLOPS + OCTAVE_FLOPS) + TERRAIN_FLOPS


def smoothing_flops(V, T, K):
    """Vertex placement (4 per vertex) and K dual/primal smoothing iterations:
    per triangle its centroid (6 adds + 3 divides) and its share of the three
    corner ring sums (3 x 3 adds), per vertex the 3 divides by the ring size."""
    return 4 * V + K * (9 * T + 3 * (3 * T) + 3 * V)


def mesh_bytes(V, T):
    """Compulsory output bytes of a mesh: 12 B positions + 1 B boundary per
    vertex, 12 B indices per triangle, 16 B of counts."""
    return 13 * V + 12 * T + 16


def frame_fractions(flops, nbytes, seconds, note=None):
    """Whole-frame work over its time against the FP32-VALU and HBM peaks:
    algorithmic FLOP (full-octave sampling and the smoothing) and compulsory
    mesh bytes per frame."""
    out = {"flops_per_step": int(flops), "bytes_per_step": int(nbytes),
           "tflops": round(flops / seconds / 1e12, 4),
           "fp32_frac": round(flops / seconds / 1e12 / PEAK_FP32_TFLOPS, 5),
           "hbm_gbs": round(nbytes / seconds / 1e9, 3),
           "hbm_frac": round(nbytes / seconds / 1e9 / PEAK_HBM_GBS, 6)}
    if note:
        out["note"] = note
    return out
