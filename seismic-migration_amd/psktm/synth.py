"""Seeded synthetic surveys for the PSTM workloads.

Conventions follow the reference: traces are trace-major [n_traces][n_samples];
offsets are half-offset vectors (scripts/ktmConvertXYtoOffMid.m:12); the image
time axis is one-way tau, so an output sampled every d seconds of two-way time
has dtau = d/2 (src/psktmCPU.c:78-79); velocities are RMS velocities in m/s.

* :func:`line_2d` - the 2-D line of 64 CMPs x 32 offsets (2,048 traces x
  1,000 samples), 25 Hz Ricker reflections from 3 flat and 2 dipping
  reflectors, Gaussian noise at 20 dB SNR.
* :func:`diffractor_3d` - the 3-D survey: uniform midpoints, offsets
  uniform 0-4000 m, uniform azimuths, v(x,y,t0) = 1500 + 600 t0 +
  200 sin(x/2000) cos(y/3000), 30 Hz Ricker point diffractors at DSR times
  plus Gaussian noise. Built with torch on any device (GPU at full size).
* :func:`offset_class_gathers` - 64 offset-class gathers with lognormal trace
  counts and square patches of 32-128 bins, for the batched launch.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED0 = 20261203


def ricker_np(t, f):
    a = (math.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


@dataclass
class Line2D:
    grid: tuple          # (nx, x0, dx, ny, y0, dy, nt, tau0, dtau)
    mid_x: np.ndarray
    mid_y: np.ndarray
    off_x: np.ndarray
    off_y: np.ndarray
    traces: np.ndarray   # [n_traces, n_samples] float32
    vel: np.ndarray      # [nt] float32, v_rms(t0 = 2 tau)
    dt: float
    seed: int


def line_2d(seed: int = SEED0 + 0, n_cmp: int = 64, n_off: int = 32, ns: int = 1000,
            dt: float = 0.004, nt: int = 1000, noise_db: float = 20.0,
            f_peak: float = 25.0) -> Line2D:
    """2-D PSTM line: CMPs every 12.5 m, full offsets 0..1550 m step 50 m."""
    rng = np.random.default_rng(seed)
    dcmp, doff = 12.5, 50.0
    cmp = np.repeat(np.arange(n_cmp), n_off)
    off_full = np.tile(np.arange(n_off) * doff, n_cmp)
    mid_x = (cmp * dcmp).astype(np.float32)
    off_x = (off_full / 2).astype(np.float32)
    zeros = np.zeros(n_cmp * n_off, np.float32)
    vel_fn = lambda t0: 1500.0 + 750.0 * t0
    xc = 0.5 * (n_cmp - 1) * dcmp
    # (t0 at the line centre [s], dip [deg], amplitude)
    events = [(0.6, 0.0, 1.0), (1.4, 0.0, -0.8), (2.4, 0.0, 0.6), (1.0, 10.0, 0.7), (1.9, 25.0, -0.5)]
    t = np.arange(ns) * dt
    sig = np.zeros((n_cmp * n_off, ns))
    m = mid_x.astype(np.float64)
    h = off_x.astype(np.float64)
    for t0c, dip, amp in events:
        v = vel_fn(t0c)
        th = math.radians(dip)
        t0 = t0c + 2.0 * math.sin(th) * (m - xc) / v           # zero-offset time
        tt = np.sqrt(t0 ** 2 + (2.0 * h * math.cos(th) / v) ** 2)  # DSR / Levin moveout
        sig += amp * ricker_np(t[None, :] - tt[:, None], f_peak)
    sigma = np.sqrt(np.mean(sig ** 2)) / (10 ** (noise_db / 20.0))
    traces = (sig + sigma * rng.standard_normal(sig.shape)).astype(np.float32)
    dtau = dt / 2 * (ns / nt) if nt != ns else dt / 2
    grid = (n_cmp, 0.0, dcmp, 1, 0.0, dcmp, nt, 0.0, dtau)
    tau = np.float32(0.0) + np.arange(nt, dtype=np.float32) * np.float32(dtau)
    vel = vel_fn(2.0 * tau.astype(np.float64)).astype(np.float32)
    return Line2D(grid, mid_x, zeros.copy(), off_x, zeros.copy(), traces, vel, dt, seed)


def velocity_3d_fn(x, y, t0):
    """v_rms(x, y, t0) = 1500 + 600 t0 + 200 sin(x/2000) cos(y/3000) (m/s)."""
    xp = np if isinstance(x, np.ndarray) else __import__("torch")
    return 1500.0 + 600.0 * t0 + 200.0 * xp.sin(x / 2000.0) * xp.cos(y / 3000.0)


@dataclass
class Survey3D:
    grid: tuple
    mid_x: object
    mid_y: object
    off_x: object
    off_y: object
    traces: object       # torch [n_traces, ns] float32
    vel: object          # torch [nx, ny, nt] float32
    dt: float
    aperture: float
    seed: int


def diffractor_3d(n_traces: int, *, nx=200, ny=200, dxy=25.0, nt=1500, dtau=0.002,
                  ns=1500, dt=0.004, n_diff=200, max_off=4000.0, f_peak=30.0,
                  noise_db=20.0, aperture=1500.0, seed=SEED0 + 1, device="cpu",
                  chunk=16384) -> Survey3D:
    """3-D point-diffractor survey on ``device`` (torch)."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    ext_x, ext_y = (nx - 1) * dxy, (ny - 1) * dxy
    mx = torch.rand(n_traces, generator=g, dtype=torch.float64) * ext_x
    my = torch.rand(n_traces, generator=g, dtype=torch.float64) * ext_y
    off = torch.rand(n_traces, generator=g, dtype=torch.float64) * max_off
    az = torch.rand(n_traces, generator=g, dtype=torch.float64) * (2 * math.pi)
    hx, hy = 0.5 * off * torch.cos(az), 0.5 * off * torch.sin(az)
    # diffractors: position, one-way time, amplitude
    dxp = torch.rand(n_diff, generator=g, dtype=torch.float64) * ext_x
    dyp = torch.rand(n_diff, generator=g, dtype=torch.float64) * ext_y
    tmax = (ns - 1) * dt
    dtau_d = 0.1 * tmax + torch.rand(n_diff, generator=g, dtype=torch.float64) * 0.3 * tmax
    damp = torch.where(torch.rand(n_diff, generator=g) < 0.5, -1.0, 1.0).to(torch.float64) * \
        (0.5 + torch.rand(n_diff, generator=g, dtype=torch.float64))
    noise_seed = int(torch.randint(0, 2**31 - 1, (1,), generator=g).item())
    dev = torch.device(device)
    mx32, my32, hx32, hy32 = (a.to(torch.float32) for a in (mx, my, hx, hy))
    vd = velocity_3d_fn(dxp, dyp, 2 * dtau_d)
    traces = diffraction_traces(mx32, my32, hx32, hy32, dxp, dyp, dtau_d, damp, vd, ns, dt,
                                f_peak, noise_db, noise_seed, dev, chunk)
    xs = torch.arange(nx, dtype=torch.float32) * np.float32(dxy)
    ys = torch.arange(ny, dtype=torch.float32) * np.float32(dxy)
    taus = torch.arange(nt, dtype=torch.float32) * np.float32(dtau)
    vel = velocity_3d_fn(xs[:, None, None].double(), ys[None, :, None].double(),
                         2.0 * taus[None, None, :].double()).to(torch.float32).to(dev).contiguous()
    grid = (nx, 0.0, dxy, ny, 0.0, dxy, nt, 0.0, dtau)
    return Survey3D(grid, mx32.to(dev), my32.to(dev), hx32.to(dev), hy32.to(dev), traces, vel,
                    dt, aperture, seed)


def diffraction_traces(mx32, my32, hx32, hy32, dxp, dyp, dtau_d, damp, vd, ns, dt, f_peak,
                       noise_db, noise_seed, dev, chunk=16384):
    """Point diffractors (Ricker wavelets at DSR times) plus Gaussian noise.

    Positions are the ones the operator sees, s = mid - off, r = mid + off,
    formed in FP64 from the float32 midpoints and half offsets; diffractor d
    sits at (dxp, dyp), one-way time dtau_d, RMS velocity vd, amplitude damp.
    """
    import torch
    n_traces = mx32.numel()
    sx = (mx32.double() - hx32.double()).to(dev)
    sy = (my32.double() - hy32.double()).to(dev)
    rx = (mx32.double() + hx32.double()).to(dev)
    ry = (my32.double() + hy32.double()).to(dev)
    dxp, dyp, dtau_d, damp, vd = (a.to(dev) for a in (dxp, dyp, dtau_d, damp, vd))
    traces = torch.empty((n_traces, ns), dtype=torch.float32, device=dev)
    half = int(math.ceil(2.5 / (math.pi * f_peak) / dt)) + 1     # wavelet support
    win = torch.arange(-half, half + 1, device=dev, dtype=torch.float64)
    pf2 = (math.pi * f_peak) ** 2
    for a in range(0, n_traces, chunk):
        b = min(n_traces, a + chunk)
        ts = torch.sqrt(dtau_d[None] ** 2 + ((sx[a:b, None] - dxp[None]) ** 2 +
                                               (sy[a:b, None] - dyp[None]) ** 2) / vd[None] ** 2)
        tr = torch.sqrt(dtau_d[None] ** 2 + ((rx[a:b, None] - dxp[None]) ** 2 +
                                               (ry[a:b, None] - dyp[None]) ** 2) / vd[None] ** 2)
        td = (ts + tr) / dt                                   # [c, n_diff] in samples
        n0 = torch.floor(td)
        idx = n0[..., None] + win                              # [c, n_diff, w]
        arg = (idx - td[..., None]) * dt
        val = damp[None, :, None] * (1 - 2 * pf2 * arg ** 2) * torch.exp(-pf2 * arg ** 2)
        ok = (idx >= 0) & (idx < ns)
        rows = torch.arange(b - a, device=dev)[:, None, None].expand_as(idx)
        flat = (rows * ns + idx.clamp(0, ns - 1).long())[ok]
        blk = torch.zeros((b - a) * ns, dtype=torch.float64, device=dev)
        blk.index_add_(0, flat, val[ok])
        traces[a:b] = blk.view(b - a, ns).to(torch.float32)
    if n_traces:
        rms = torch.sqrt(torch.mean(traces.double() ** 2)).item()
        sigma = rms / (10 ** (noise_db / 20.0))
        ng = torch.Generator(device=dev).manual_seed(noise_seed)
        for a in range(0, n_traces, chunk * 8):
            b = min(n_traces, a + chunk * 8)
            traces[a:b] += sigma * torch.randn((b - a, ns), generator=ng, device=dev,
                                               dtype=torch.float32)
    return traces


@dataclass
class Gathers:
    grids: list          # per gather (nx, x0, dx, ny, y0, dy, nt, tau0, dtau)
    first: list          # first trace of each gather in the concatenated arrays
    counts: list
    mid_x: object
    mid_y: object
    off_x: object
    off_y: object
    traces: object       # torch [sum counts, ns] float32
    vel: object          # torch [nt] float32, v_rms(t0 = 2 tau)
    dt: float
    aperture: float
    seed: int
    params: dict


def offset_class_gathers(n_gathers: int = 64, *, mu=math.log(11000.0), sigma=0.8,
                         n_min=2000, n_max=60000, side_min=32, side_max=128, dxy=25.0,
                         nt=1000, dtau=0.002, ns=1000, dt=0.004, class_width=62.5,
                         aperture=1500.0, n_diff=4, f_peak=30.0, noise_db=20.0,
                         count_scale=1.0, seed=SEED0 + 3, device="cpu") -> Gathers:
    """Offset-class gathers: gather g holds traces whose full offset lies in
    [class_width*g, class_width*(g+1)); trace counts lognormal(mu, sigma)
    clipped to [n_min, n_max] (times count_scale); square patches with a side
    uniform in [side_min, side_max] bins; midpoints uniform over the patch;
    v_rms(t0) = 1500 + 600 t0 shared by all gathers."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    grids, first, counts = [], [], []
    mids_x, mids_y, offs_x, offs_y, trs = [], [], [], [], []
    vel_fn = lambda t0: 1500.0 + 600.0 * t0
    total = 0
    for k in range(n_gathers):
        n = float(torch.exp(mu + sigma * torch.randn(1, generator=g, dtype=torch.float64)))
        n = int(max(1, round(min(max(n, n_min), n_max) * count_scale)))
        side = int(torch.randint(side_min, side_max + 1, (1,), generator=g))
        ox = float(torch.randint(0, 400, (1,), generator=g)) * dxy
        oy = float(torch.randint(0, 400, (1,), generator=g)) * dxy
        ext = (side - 1) * dxy
        mx = ox + torch.rand(n, generator=g, dtype=torch.float64) * ext
        my = oy + torch.rand(n, generator=g, dtype=torch.float64) * ext
        off = (k + torch.rand(n, generator=g, dtype=torch.float64)) * class_width
        az = torch.rand(n, generator=g, dtype=torch.float64) * (2 * math.pi)
        hx, hy = 0.5 * off * torch.cos(az), 0.5 * off * torch.sin(az)
        dxp = ox + torch.rand(n_diff, generator=g, dtype=torch.float64) * ext
        dyp = oy + torch.rand(n_diff, generator=g, dtype=torch.float64) * ext
        tmax = (ns - 1) * dt
        dtd = 0.1 * tmax + torch.rand(n_diff, generator=g, dtype=torch.float64) * 0.3 * tmax
        damp = torch.where(torch.rand(n_diff, generator=g) < 0.5, -1.0, 1.0).to(torch.float64)
        noise_seed = int(torch.randint(0, 2**31 - 1, (1,), generator=g))
        m32 = [a.to(torch.float32) for a in (mx, my, hx, hy)]
        tr = diffraction_traces(*m32, dxp, dyp, dtd, damp, vel_fn(2 * dtd), ns, dt, f_peak,
                                noise_db, noise_seed, torch.device(device))
        grids.append((side, ox, dxy, side, oy, dxy, nt, 0.0, dtau))
        first.append(total)
        counts.append(n)
        total += n
        mids_x.append(m32[0]); mids_y.append(m32[1]); offs_x.append(m32[2]); offs_y.append(m32[3])
        trs.append(tr)
    dev = torch.device(device)
    cat = lambda xs: torch.cat(xs).to(dev).contiguous()
    tau = torch.arange(nt, dtype=torch.float32) * np.float32(dtau)
    vel = vel_fn(2.0 * tau.double()).to(torch.float32).to(dev)
    return Gathers(grids, first, counts, cat(mids_x), cat(mids_y), cat(offs_x), cat(offs_y),
                   torch.cat(trs).contiguous(), vel, dt, aperture, seed,
                   dict(mu=mu, sigma=sigma, n_min=n_min, n_max=n_max, side_min=side_min,
                        side_max=side_max, class_width=class_width, count_scale=count_scale))
