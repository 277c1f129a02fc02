"""A copy of known size through the library (y = 1 * x, mc_wfn_axpby with
beta = 0: reads N bytes, writes N bytes), run under rocprofv3 --pmc to scale
the FETCH_SIZE / WRITE_SIZE counters of the measured kernels:

    rocprofv3 --pmc FETCH_SIZE -d <out> -o copy -- python3 -m cpmd_amd.profile_copy
    rocprofv3 --pmc WRITE_SIZE -d <out> -o copy -- python3 -m cpmd_amd.profile_copy

The kernel is mc::k_axpby; it prints the byte count it moved.
"""
from __future__ import annotations

import argparse
import sys

from . import ops
from . import system as S


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024, help="bytes read (and written) per copy, MiB")
    ap.add_argument("--repeat", type=int, default=3)
    args = ap.parse_args()
    sysm = S.make_system(12.0, 8, 40.0, nstate=16)
    ctx = ops.Context(sysm.cell, sysm.gset)
    per_state = 16 * ctx.ld
    nstate = max(1, (args.mib << 20) // per_state)
    x = ctx.alloc_wfn(nstate)
    y = ctx.alloc_wfn(nstate)
    for _ in range(args.repeat):
        ctx.axpby_wfn(1.0, x, 0.0, y, nstate)
    ctx.synchronize()
    print(f"copy: {args.repeat} x {nstate * per_state} bytes read and written by mc::k_axpby", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
