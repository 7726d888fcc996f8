"""Known-traffic copy kernel for scaling hardware byte counters.

`python -m parallel_anns_amd.calibrate --bytes N` copies N bytes between two
HBM buffers with a plain streaming kernel (N read + N written) and prints the
device time.  Profiled with `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE),
the counter reading divided by N gives the factor that turns the same counter
on the search kernel into bytes.
"""
import argparse
import json

from .search import Index


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 30)
    args = ap.parse_args()
    idx = Index(0)
    idx.set_base(__import__("numpy").zeros((8, 96), "float32"))
    ms = idx.device_copy(args.bytes)
    print(json.dumps({"copy_bytes": args.bytes, "ms": ms, "GBps_read_plus_write": 2 * args.bytes / ms / 1e6}))
    idx.close()


if __name__ == "__main__":
    main()
