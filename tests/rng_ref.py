"""CPU Philox4x32-10 and the Box-Muller mapping the HIP kernels use.

Test infrastructure: checks the device generator (rp_sample_deviates)."""
import numpy as np

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ c3 ^ k1) & MASK, p0 & MASK
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c0, c1, c2, c3


def normals(seed, pid, step, ndev):
    out = []
    blk = 0
    while len(out) < ndev:
        ctr = (blk, pid & MASK, (pid >> 32) & MASK, step & MASK)
        key = (seed & MASK, ((seed >> 32) ^ (step >> 32)) & MASK)
        r = philox4x32_10(ctr, key)
        u = [(v + 0.5) * 2.0 ** -32 for v in r]
        ra = np.sqrt(-2.0 * np.log(u[0]))
        rb = np.sqrt(-2.0 * np.log(u[2]))
        tp = 6.283185307179586
        out += [ra * np.cos(tp * u[1]), ra * np.sin(tp * u[1]), rb * np.cos(tp * u[3]), rb * np.sin(tp * u[3])]
        blk += 1
    return np.array(out[:ndev])
