"""Calibration workload for HBM byte counters: a plain device copy of a known
size (1 GiB read + 1 GiB written per dispatch), run under
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` beside the renderer so that
the counters' KiB per dispatch can be scaled to bytes.

    python3 livre_amd/hbm_probe.py [--gib 1] [--reps 4]
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=int, default=1)
    ap.add_argument("--reps", type=int, default=4)
    a = ap.parse_args()
    n = a.gib << 30
    src = torch.empty(n, dtype=torch.uint8, device="cuda").random_(0, 255)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copied %d x %d bytes" % (a.reps, n), flush=True)


if __name__ == "__main__":
    main()
