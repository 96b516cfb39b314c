"""Seeded inputs shared by the benchmark and the tests.

- `xavier_blob`: relaxation-time network 6 -> 32 -> 32 -> 1 with
  Xavier-uniform weights (numpy PCG64 seeded; W1 and W2 rounded to bf16
  values by default), in the parameter layout of
  ippi_set_mlp (W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3, mu[6],
  1/sigma[6]);
- `place_cylinders`: cylinders with diameters uniform in [dmin, dmax] at
  non-overlapping positions by rejection sampling, each marker support
  (4x4 delta stencil) kept `gap` lattice units clear of every other body's
  support and of the domain edges;
- `markers_for`: marker count for ~1 lu spacing on the circle;
- `init_blob` / `bf16_weights`: the starting point of training on solver
  pairs, and W1 / W2 rounded to bf16 after it.

The reference's own cylinder markers follow mesh.cpp:180-181; its network
is a libtorch model loaded per call (mesh.cpp:2767-3260) and has no
counterpart here beyond the parameter layout above.
"""
import math

import numpy as np

MLP_PARAMS = 6 * 32 + 32 + 32 * 32 + 32 + 32 + 1
MLP_BLOB = MLP_PARAMS + 12


def round_bf16(x):
    """Nearest bf16 values (ties to even) of x, as float64."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


def xavier_blob(seed, b3=-5.0, mu=(1.0, 0.05, 0.0, 2e-3, 0.0, 0.0), sig=(2e-3, 0.02, 0.02, 2e-3, 5e-4, 5e-4),
                w3_scale=0.3, bf16_weights=True):
    """Seeded Xavier-uniform network. With the default b3 the output
    softplus(.) + 0.5 sits a few 1e-3 above 0.5, so the >= 0.501 guard
    accepts it at typical nodes. bf16_weights: W1 and W2 rounded to bf16
    values, which the FP32 path multiplies exactly (two MFMAs per product
    instead of three)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_BLOB)

    def xav(n_in, n_out):
        a = math.sqrt(6.0 / (n_in + n_out))
        return rng.uniform(-a, a, (n_out, n_in))
    p[0:192] = xav(6, 32).ravel()
    p[224:1248] = xav(32, 32).ravel()
    if bf16_weights:
        p[0:192] = round_bf16(p[0:192])
        p[224:1248] = round_bf16(p[224:1248])
    p[1280:1312] = xav(32, 1).ravel() * w3_scale
    p[1312] = b3
    p[1313:1319] = mu
    p[1319:1325] = 1.0 / np.asarray(sig, dtype=np.float64)
    return p


def init_blob(x, omega, seed):
    """Starting point for training on features x [n, 6] and labels omega
    [n]: Xavier-uniform W1, W2, W3 (numpy PCG64 seeded), mu / 1/sigma from
    the features, and b3 placing softplus(b3) + 0.5 at the median label."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_BLOB)
    a1, a2 = math.sqrt(6 / 38), math.sqrt(6 / 64)
    p[0:192] = rng.uniform(-a1, a1, 192)
    p[224:1248] = rng.uniform(-a2, a2, 1024)
    p[1280:1312] = rng.uniform(-0.3, 0.3, 32)
    tau = 1.0 / float(np.median(omega))
    p[1312] = math.log(math.expm1(max(tau - 0.5, 1e-6)))
    p[1313:1319] = x.mean(0)
    p[1319:1325] = 1.0 / np.maximum(x.std(0), 1e-30)
    return p


def bf16_weights(blob):
    """The blob with W1 and W2 rounded to bf16 values (the FP32 path then
    multiplies them exactly, two MFMAs per product)."""
    b = np.array(blob, dtype=np.float64)
    b[0:192] = round_bf16(b[0:192])
    b[224:1248] = round_bf16(b[224:1248])
    return b


BNN_PARAMS = 15 * 128 + 128 + 6 * (128 * 128 + 128) + 128 + 1


def boundary_blob(seed, out_bias=0.55, out_scale=0.02):
    """Seeded boundary-row / column network (the reference's AutoT0x shape:
    Linear(15,128), 6 x Linear(128,128), Linear(128,1)), Xavier-uniform
    hidden weights; the output sits near out_bias, above the 0.501 guard."""
    rng = np.random.default_rng(seed)
    p = []

    def lin(nin, nout):
        a = math.sqrt(6.0 / (nin + nout))
        p.append(rng.uniform(-a, a, (nout, nin)).ravel())
        p.append(rng.uniform(-0.1, 0.1, nout))
    lin(15, 128)
    for _ in range(6):
        lin(128, 128)
    p.append(rng.uniform(-1, 1, 128) * out_scale)
    p.append([out_bias])
    return np.concatenate(p)


def markers_for(diameter):
    return int(round(math.pi * diameter))


def place_cylinders(nx, ny, n, dmin, dmax, seed, gap=4.0, max_tries=100000):
    """[(cx, cy, D, M)] for n cylinders; raises if they do not fit."""
    rng = np.random.default_rng(seed)
    out = []
    reach = 2.0  # delta support half-width
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place the cylinders")
        d = float(rng.uniform(dmin, dmax))
        r = 0.5 * d + reach + 1.0
        cx = float(rng.uniform(r + 1, nx - r - 2))
        cy = float(rng.uniform(r + 1, ny - r - 2))
        ok = True
        for (ox, oy, od, _) in out:
            # bounding boxes of the supports must be gap apart (patches may not overlap)
            rr = 0.5 * od + reach + 1.0
            if abs(cx - ox) < r + rr + gap and abs(cy - oy) < r + rr + gap:
                ok = False
                break
        if ok:
            out.append((cx, cy, d, markers_for(d)))
    return out
