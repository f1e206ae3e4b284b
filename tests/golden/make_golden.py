"""Writes the golden fixtures of this directory from mathematics alone.

The reference (wjb19/seismic-migration) ships no tests, fixtures or golden
vectors, and its C sources need struct and mask definitions it never ships,
so no vector can come from running it. These fixtures are therefore computed
here in FP64 numpy straight from the definitions - explicit source and
receiver positions, the double-square-root time, the weight
t*sqrt(tau^2/t) (src/psktmCPU.c:78-85), the rho filter's closed forms on
cosines (scripts/ktmApplyRhoFilter.m:31-58) and the IBM single-precision
format (src/ioFunctions.c:4-40) - never from the oracle's code, and then
frozen, so that the oracle and the GPU kernels are both held to them.

Run from the repository root: python tests/golden/make_golden.py
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def axis32(start, step, n):
    """Binary32 axis values start + (float)i * step (src/psktmFunctions.c:28-34)."""
    i = np.arange(n, dtype=np.float32)
    return (np.float32(start) + i * np.float32(step)).astype(np.float32)


def migration_case():
    """A small 3-D survey and its FP64 LINEAR image, per-update weight and
    gained form (weight tau; traces times sqrt(n dt)), without aperture."""
    rng = np.random.default_rng(20261204)
    nx, ny, nt, ns = 6, 5, 90, 150
    x0, dx, y0, dy, tau0, dtau = 0.0, 30.0, 10.0, 25.0, 0.0, 0.004
    dt = float(np.float32(0.008))       # the sample interval as the API carries it
    ntr = 24
    mid = rng.uniform(-60, 200, (ntr, 2)).astype(np.float32)
    off = rng.uniform(-250, 250, (ntr, 2)).astype(np.float32)
    n = np.arange(ns)
    f = rng.uniform(0.02, 0.07, (ntr, 1))
    ph = rng.uniform(0, 2 * np.pi, (ntr, 1))
    traces = (np.cos(2 * np.pi * f * n[None] + ph) * rng.uniform(0.5, 2.0, (ntr, 1))).astype(np.float32)
    vel = (1600 + 900 * np.linspace(0, 1, nt)).astype(np.float32)
    xx, yy, zz = axis32(x0, dx, nx), axis32(y0, dy, ny), axis32(tau0, dtau, nt)
    src = mid.astype(np.float64) - off.astype(np.float64)
    rec = mid.astype(np.float64) + off.astype(np.float64)
    X, Y = np.meshgrid(xx.astype(np.float64), yy.astype(np.float64), indexing="ij")
    tau = zz.astype(np.float64)
    v = vel.astype(np.float64)
    img = np.zeros((nx, ny, nt))
    img_g = np.zeros((nx, ny, nt))
    gain = np.sqrt(n * dt)
    for k in range(ntr):
        ds2 = (src[k, 0] - X) ** 2 + (src[k, 1] - Y) ** 2
        dr2 = (rec[k, 0] - X) ** 2 + (rec[k, 1] - Y) ** 2
        t = (np.sqrt(tau[None, None] ** 2 + ds2[..., None] / v ** 2)
             + np.sqrt(tau[None, None] ** 2 + dr2[..., None] / v ** 2))
        q = t / dt
        m = np.floor(q).astype(np.int64)
        valid = (m >= 0) & (m + 1 < ns)
        mm = np.where(valid, m, 0)
        fr = q - m
        d = traces[k].astype(np.float64)
        g = d * gain
        lerp = (1 - fr) * d[mm] + fr * d[mm + 1]
        lerp_g = (1 - fr) * g[mm] + fr * g[mm + 1]
        img += np.where(valid, lerp * tau * np.sqrt(t), 0.0)
        img_g += np.where(valid, lerp_g * tau, 0.0)
    np.savez(os.path.join(HERE, "migration_3d_small.npz"),
             grid=np.array([nx, ny, nt], np.int32),
             axes=np.array([x0, dx, y0, dy, tau0, dtau], np.float32),
             dt=np.float32(dt), mid=mid, off=off, traces=traces, vel=vel,
             image_linear=img, image_linear_gained=img_g)


def rho_case():
    """Rho filter outputs for cosines x[n] = cos(2 pi k0 n / M), m = M: with
    w0 = 2 pi k0 / (M dr) the reference's filter (no 1/M in the forward, 1/M
    in the inverse, frequencies w = 2 pi k / dr) returns opt 1: M w0 cos,
    opt 2: sqrt(M w0) cos(. + pi/4), opt 3: -M w0 sin."""
    M, dw_us = 256, 4000.0
    dr = dw_us * 1e-6
    n = np.arange(M)
    # an even trace count: the reference's 2-D fftshift rotates odd counts
    # by one row (scripts/ktmApplyRhoFilter.m), which this fixture avoids
    ks = np.array([1, 5, 17, 100])
    x = np.cos(2 * np.pi * ks[:, None] * n[None] / M)
    out = {}
    for k0 in ks:
        w0 = 2 * np.pi * k0 / (M * dr)
        ph = 2 * np.pi * k0 * n / M
        out.setdefault(1, []).append(M * w0 * np.cos(ph))
        out.setdefault(2, []).append(math.sqrt(M * w0) * np.cos(ph + np.pi / 4))
        out.setdefault(3, []).append(-M * w0 * np.sin(ph))
    np.savez(os.path.join(HERE, "rho_cosines.npz"), x=x.astype(np.float32),
             dw_us=np.float64(dw_us), opt1=np.array(out[1]), opt2=np.array(out[2]),
             opt3=np.array(out[3]))


def ibm_case():
    """IBM System/360 singles: sign, base-16 exponent biased by 64, 24-bit
    fraction. Exact value (-1)^s * frac / 2^24 * 16^(e - 64) (the format's
    definition) and the reference's formula, which divides by 2^24 - 1."""
    words = [0x41100000, 0x42640000, 0xC276A000, 0x00000000, 0x40800000,
             0xC1100000, 0x3F100000, 0x4110000F, 0x45123456, 0xBE7FFFFF]
    exact, ref = [], []
    for w in words:
        s = -1.0 if w & 0x80000000 else 1.0
        e = (w >> 24) & 0x7F
        fr = w & 0x00FFFFFF
        exact.append(s * fr / float(1 << 24) * 16.0 ** (e - 64))
        ref.append(s * (fr / float((1 << 24) - 1)) * 16.0 ** (e - 64))
    np.savez(os.path.join(HERE, "ibm_words.npz"), words=np.array(words, np.uint32),
             exact=np.array(exact), reference_formula=np.array(ref))


if __name__ == "__main__":
    migration_case()
    rho_case()
    ibm_case()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
