"""bench.py reads the HBM traffic of the launch shape the autotuner chose
from the committed rocprofv3 summaries under profiles/; every one-iteration
shape the autotuner can choose must be found there, or roofline.traffic
would be reported as null."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class _Plan:
    def __init__(self, tile, threads, ipl=1):
        self.tile_verts, self.block_threads, self.iters_per_launch = tile, threads, ipl


@pytest.mark.parametrize("path,algorithmic", [
    (bench.TRAFFIC_FILE, bench.algorithmic_bytes(65536, 130050, 194565)),
    (bench.TRAFFIC_FILE_1024, bench.algorithmic_bytes(1048576, 2093058, 3137538)),
    (bench.TRAFFIC_FILE_BATCH, bench.algorithmic_bytes(6586714, 12986844, 19387486)),
])
def test_every_shape_has_profiled_traffic(path, algorithmic):
    for tile in (32, 64, 128, 256):
        for threads in (256, 512, 1024):
            t = bench.profiled_traffic(path, _Plan(tile, threads))
            assert t is not None, (os.path.basename(path), tile, threads)
            # real traffic lies above the algorithmic bytes and below twice them
            assert algorithmic < t < 2 * algorithmic, (tile, threads, t, algorithmic)
    assert bench.profiled_traffic(path, _Plan(128, 256, 2)) is None
