"""The serving configuration: 32 concurrent sessions over the five volumes
(256^3 and 512^3 uint8 blobs, 1024^3 and 2048^3 uint16 noise, 1536^3 float32
sines), viewports from 720p to 4K, seeded random-walk cameras and TFs, served
by the continuous-batching loop under a Poisson burst of 200 requests/s for
one second. Every request is rendered or superseded by a newer one of its
session (none fails), and the last served frame of sessions 0, 7, 15 and 31
equals the same frame rendered alone by lvr_render bit for bit (the alone
frames are the single-frame path that the oracle tests check)."""
import argparse
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def test_c5_poisson_burst(lvr):
    args = argparse.Namespace(max_batch=2, pool_frac=1.0)
    res = bench.serve_measure(args, (200,), 1.0)
    row = res["rates"][0]
    print("c5:", row["requests"], "requests,", row["rendered"], "rendered,", row["superseded"],
          "superseded, p50 %.1f ms p99 %.1f ms" % (row["p50_ms"], row["p99_ms"]), flush=True)
    assert row["failed"] == 0 and row["rendered"] > 0
    assert row["rendered"] + row["superseded"] == row["requests"]
    assert res["served_equals_alone"]
