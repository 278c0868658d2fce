"""Write tests/golden/truth_errors.npz: the FP64 error floors of the
compiled reference and of the FP64 oracle against the binary128 truth, per
item, on the accuracy samples of the benchmark workloads.

Runs in the build container only (it needs oracle/_ref/libsolidfmm_ref.so,
compiled from the reference's own sources by oracle/Makefile).  Inputs are
regenerated from their seeds (solidfmm_amd.workload), so the file holds only
per-item numbers:
  e_ref, e_orc      max over degrees of the per-degree normwise relative
                    error of the reference (single-thread m2m/m2l/l2l,
                    translations.hpp:122-170, in calls of <= 12 items to stay
                    clear of the AVX-512 lane defect) and of the oracle,
                    against the truth's naive O(P^4) route;
  u_ref, u_orc      the same errors in units of the rounding scale
                    u ||M_n|| (oracle.error_scale), max over degrees.
The GPU tests recompute the truth by the rotation route and must find the
same e_ref, which ties their truth to this file's naive one.

    python tests/golden/gen_truth.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from solidfmm_amd import workload as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truth_errors.npz")

# name: (op, p_in, p_out, count, seed, shift) -- square samples; "mixed": the
# mixed-order M2L workload's generator.
SAMPLES = {
    "m2l_p10": (W.M2L, 10, 10, 4096, 0x5011DF00, "m2l"),   # the headline batch's first 4,096 items
    "m2m_p16": (W.M2M, 16, 16, 2000, 0x5011DF00, "octant"),
    "l2l_p16": (W.L2L, 16, 16, 2000, 0x5011DF00, "octant"),
    "m2l_p16": (W.M2L, 16, 16, 2000, 0x5011DF00, "m2l"),
    "m2l_p24": (W.M2L, 24, 24, 2000, 0x5011DF00, "m2l"),
    "mixed": (W.M2L, 4, 20, 2000, 0x5011DF02, "mixed"),
}


def sample(name):
    """(op, items, src, dst) of a sample, zero targets."""
    op, a, b, n, seed, shift = SAMPLES[name]
    if shift == "mixed":
        items, src, dst = W.mixed_batch(n, seed, a, b, op=op)
    else:
        items, src, dst = W.uniform_batch(op, a, b, n, seed, shift=shift)
    return op, items, src, dst


def errors(got, hi, lo, mag, items):
    """Per item: (max relative error, max error in rounding units)."""
    e = np.zeros(len(items))
    u = np.zeros(len(items))
    for po in np.unique(items["p_out"]):
        po = int(po)
        idx = np.nonzero(items["p_out"] == po)[0]
        cols = items["dst"][idx].astype(np.int64)[:, None] + np.arange(po * (po + 1))
        ed = oracle.truth_degree_errors(got[cols], hi[cols], lo[cols], po)
        s = oracle.error_scale(mag[cols], hi[cols], po)
        e[idx] = ed.max(axis=1)
        u[idx] = np.where(s > 0, ed / np.where(s > 0, s, 1.0), 0.0).max(axis=1)
    return e, u


def reference_output(ref, op, items, src, dst):
    out = np.zeros_like(dst)
    P = max(int(items["p_in"].max()), int(items["p_out"].max()))
    ref.translate(op, P, items, src, out)
    return out


def main():
    T = oracle.Truth()
    O = oracle.Oracle()
    R = oracle.Reference()
    arrays = {}
    for name in SAMPLES:
        op, items, src, dst = sample(name)
        hi, lo = T.translate(op, items, src, len(dst), method=T.NAIVE)
        _, _, mag = T.translate(op, items, src, len(dst), magnitude=True)
        r = reference_output(R, op, items, src, dst)
        o = O.translate(op, items, src, np.zeros_like(dst))
        for who, v in (("ref", r), ("orc", o)):
            e, u = errors(v, hi, lo, mag, items)
            arrays[f"{name}/e_{who}"] = e
            arrays[f"{name}/u_{who}"] = u
        print(f"{name}: ref p99 {np.percentile(arrays[name + '/e_ref'], 99):.2e} max {arrays[name + '/e_ref'].max():.2e} "
              f"units max {arrays[name + '/u_ref'].max():.1f}", flush=True)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
