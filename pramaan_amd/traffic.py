"""Dispatches for measuring the HBM traffic of the batch iterate kernels with
rocprofv3 counters: the two iterate kernels on the 500-window rolling-horizon
batch with every window running (pm_batch_time_kernel), then a plain copy of
a known size (256 MiB) that calibrates the counters.

    rocprofv3 --pmc FETCH_SIZE --kernel-trace -d <dir> -o f -- python3 -m pramaan_amd.traffic
    rocprofv3 --pmc WRITE_SIZE --kernel-trace -d <dir> -o w -- python3 -m pramaan_amd.traffic

Per dispatch, bytes = counter (KiB) x 1024 x (copy bytes / the same counter
on the copy kernel).
"""
from __future__ import annotations

import argparse


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--windows", type=int, default=500)
    ap.add_argument("--reps", type=int, default=4)
    a = ap.parse_args(argv)
    import torch

    from . import _capi as P
    rh = P.RollingHorizon()
    lps = [rh.window(w) for w in range(min(a.windows, rh.windows))]
    with P.Batch(lps, tolerance=1e-6, check_frequency=64) as b:
        print("iterate kernels (A, A^T):", b.spmv_variants(), flush=True)
        for which, name in ((P.PM_KERNEL_SPMV_AT, "A^T y + primal"), (P.PM_KERNEL_SPMV_A, "A xe + dual")):
            ms, by = b.time_kernel(which, a.reps)
            print(f"{name}: {ms * 1e3:.1f} us, algorithmic {by / 1e6:.1f} MB", flush=True)
    src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy: {src.numel() * 4} bytes", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
