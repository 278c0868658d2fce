"""Seeded synthetic workloads for the translation operators.

Random numbers come from a counter-based SplitMix64 stream: output ``i``
(1-based) of stream ``s`` under seed ``seed`` is ``mix(base + i*GAMMA)`` with
``base = seed ^ (s * 0xD1B54A32D192ED03)`` and ``mix`` the SplitMix64
finaliser; a uniform in [0,1) takes the top 53 bits.  Every quantity draws
from its own stream, so inputs do not depend on generation order.

Coefficients follow the solid layout of the reference (solid.hpp:59-60):
``Re C_n^m`` at ``n(n+1)+m`` and ``Im C_n^m`` at ``(n+1)^2+m``, P(P+1) reals
for order P (degrees 0..P-1).  ``Im C_n^0`` is zero, as the folded swap
matrices ignore it (operator_data.cpp:137-143).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
STREAM_MUL = np.uint64(0xD1B54A32D192ED03)

# One translation request: offsets are in reals into the source and target
# pools; the shift is new centre minus old centre.
ITEM_DTYPE = np.dtype([
    ("src", "<u8"), ("dst", "<u8"), ("p_in", "<u4"), ("p_out", "<u4"),
    ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
])
assert ITEM_DTYPE.itemsize == 48

M2M, M2L, L2L = 0, 1, 2
OP_NAMES = {M2M: "m2m", M2L: "m2l", L2L: "l2l"}


def solid_size(P):
    """Reals per expansion of order P (degrees 0..P-1)."""
    return P * (P + 1)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed, stream, n, offset=0):
    """``n`` uniforms in [0,1) from stream ``stream``, outputs offset+1..offset+n."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream) * STREAM_MUL)
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = _mix(base + i * GAMMA)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(seed, stream, n, lo, hi, offset=0):
    return lo + (hi - lo) * uniform01(seed, stream, n, offset)


def imag_m0_mask(P):
    """Boolean mask of the Im C_n^0 slots of an order-P solid."""
    mask = np.zeros(solid_size(P), dtype=bool)
    for n in range(P):
        mask[(n + 1) * (n + 1)] = True
    return mask


def random_solids(seed, stream, count, P):
    """``count`` order-P solids, Re and Im i.i.d. U[-1,1], Im C_n^0 = 0."""
    s = solid_size(P)
    a = uniform(seed, stream, count * s, -1.0, 1.0).reshape(count, s)
    a[:, imag_m0_mask(P)] = 0.0
    return a


def random_directions(seed, stream, count):
    """Uniform unit vectors: z ~ U[-1,1], phi ~ U[0,2pi)."""
    z = uniform(seed, stream, count, -1.0, 1.0)
    phi = uniform(seed, stream + 1, count, 0.0, 2.0 * np.pi)
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


def random_shifts(seed, stream, count, rmin, rmax):
    """Uniform directions with |shift| ~ U[rmin, rmax]."""
    d = random_directions(seed, stream, count)
    r = uniform(seed, stream + 2, count, rmin, rmax)
    return d * r[:, None]


def octant_shifts(seed, stream, count):
    """Child-to-parent shifts: (+-0.5, +-0.5, +-0.5) * U[0.5, 1] with random signs."""
    s = uniform(seed, stream, count, 0.5, 1.0)
    signs = np.where(uniform01(seed, stream + 1, 3 * count).reshape(count, 3) < 0.5, -1.0, 1.0)
    return 0.5 * signs * s[:, None]


def make_items(src_off, dst_off, p_in, p_out, shifts):
    n = len(src_off)
    it = np.zeros(n, dtype=ITEM_DTYPE)
    it["src"] = src_off
    it["dst"] = dst_off
    it["p_in"] = p_in
    it["p_out"] = p_out
    it["x"] = shifts[:, 0]
    it["y"] = shifts[:, 1]
    it["z"] = shifts[:, 2]
    return it


def uniform_batch(op, P_in, P_out, count, seed, shift="m2l", distinct_targets=True, n_targets=None):
    """A batch of ``count`` translations with fixed orders.

    Returns (items, src_pool, dst_pool): sources are random solids packed
    back to back, targets start at zero.  ``shift`` is "m2l" (|x| in [2,4]) or
    "octant".  With ``distinct_targets`` each item writes its own target;
    otherwise items cycle over ``n_targets`` targets.
    """
    ss, so = solid_size(P_in), solid_size(P_out)
    src = random_solids(seed, 1, count, P_in).reshape(-1)
    nt = count if distinct_targets else int(n_targets)
    dst = np.zeros(nt * so)
    if shift == "m2l":
        sh = random_shifts(seed, 10, count, 2.0, 4.0)
    elif shift == "octant":
        sh = octant_shifts(seed, 20, count)
    else:
        raise ValueError(shift)
    src_off = np.arange(count, dtype=np.uint64) * np.uint64(ss)
    dst_off = (np.arange(count, dtype=np.uint64) % np.uint64(nt)) * np.uint64(so)
    items = make_items(src_off, dst_off, P_in, P_out, sh)
    return items, src, dst


def mixed_batch(count, seed, pmin=4, pmax=20, op=M2L):
    """Batch with P_in, P_out i.i.d. uniform on {pmin..pmax}, one target each,
    blocks packed back to back.  Shifts as in uniform_batch: |x| in [2,4]
    for M2L, octant offsets for M2M and L2L."""
    span = pmax - pmin + 1
    pin = pmin + np.minimum((uniform01(seed, 30, count) * span).astype(np.int64), span - 1)
    pout = pmin + np.minimum((uniform01(seed, 31, count) * span).astype(np.int64), span - 1)
    sin = pin * (pin + 1)
    sout = pout * (pout + 1)
    src_off = np.concatenate([[0], np.cumsum(sin)[:-1]]).astype(np.uint64)
    dst_off = np.concatenate([[0], np.cumsum(sout)[:-1]]).astype(np.uint64)
    src = uniform(seed, 32, int(sin.sum()), -1.0, 1.0)
    for n in range(pmax):  # Im C_n^0 = 0 in every block that has degree n
        has = pin > n
        src[src_off[has].astype(np.int64) + (n + 1) * (n + 1)] = 0.0
    dst = np.zeros(int(sout.sum()))
    sh = random_shifts(seed, 40, count, 2.0, 4.0) if op == M2L else octant_shifts(seed, 41, count)
    items = make_items(src_off, dst_off, pin.astype(np.uint32), pout.astype(np.uint32), sh)
    return items, src, dst
