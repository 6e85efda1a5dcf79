"""Seeded synthetic inputs for benchmarks, serving demos and tests.

Gaussian fields with P(k) = A (k / k0)^n (A = 2e4 (Mpc/h)^3, k0 = 0.1 h/Mpc,
n = -1.5 by default), a local quadratic transform
delta -> delta + f (delta^2 - <delta^2>), Zel'dovich-displaced lattice
particles, and correlated second fields for cross spectra.

Convention: white noise w(x) ~ N(0,1) from numpy's PCG64(seed); its r2c
transform scaled by 1/N^3 is multiplied by sqrt(P(k) N^3 / V), so that
<|delta(k)|^2> = P(k)/V in the 1/N^3 convention of the estimator, with
P(0) = 0.  The transforms use scipy.fft with all host cores.
"""
import os

import numpy as np
import scipy.fft as sfft

P_AMP = 2.0e4
P_K0 = 0.1
P_INDEX = -1.5


def _workers():
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def _kgrid(N, box):
    kf = 2.0 * np.pi / box
    f = np.fft.fftfreq(N, 1.0 / N)
    fz = np.arange(N // 2 + 1)
    return kf, f, fz


def power_law(k, amp=P_AMP, k0=P_K0, index=P_INDEX):
    k = np.asarray(k, dtype=np.float64)
    with np.errstate(divide='ignore'):
        p = amp * (k / k0) ** index
    return np.where(k > 0, p, 0.0)


def gaussian_spectrum(N, box, seed, amp=P_AMP, k0=P_K0, index=P_INDEX):
    """Half-complex delta(k) (1/N^3 convention) of a seeded Gaussian field."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((N, N, N))
    wk = sfft.rfftn(w, workers=_workers()) / float(N) ** 3
    del w
    kf, f, fz = _kgrid(N, box)
    vol = float(box) ** 3
    for ix in range(N):  # plane by plane keeps the temporaries small
        k = kf * np.sqrt(f[ix] ** 2 + f[:, None] ** 2 + fz[None, :] ** 2)
        wk[ix] *= np.sqrt(power_law(k, amp, k0, index) * float(N) ** 3 / vol)
    return wk


def field_from_spectrum(dk, N):
    return sfft.irfftn(dk, s=(N, N, N), workers=_workers()) * float(N) ** 3


def gaussian_field(N, box, seed, amp=P_AMP, k0=P_K0, index=P_INDEX):
    return field_from_spectrum(gaussian_spectrum(N, box, seed, amp, k0, index), N)


def local_quadratic(delta, f_nl=0.5):
    """delta + f_nl (delta^2 - <delta^2>_grid), in place on a float64 copy."""
    d = np.array(delta, dtype=np.float64)
    d2 = d * d
    d += f_nl * (d2 - d2.mean())
    return d


def quadratic_field(N, box, seed, dtype=np.float32, f_nl=0.5):
    """The seeded local-quadratic test field."""
    return local_quadratic(gaussian_field(N, box, seed), f_nl).astype(dtype)


def correlated_quadratic_pair(N, box, seed_a, seed_c, rho=0.8, dtype=np.float32, f_nl=0.5):
    """(A, B) with B built from rho*G_A + sqrt(1-rho^2)*G_C, both quadratic.

    The Gaussian parts of A and B have correlation coefficient rho.
    """
    ga = gaussian_field(N, box, seed_a)
    gc = gaussian_field(N, box, seed_c)
    gb = rho * ga + np.sqrt(1.0 - rho * rho) * gc
    del gc
    a = local_quadratic(ga, f_nl).astype(dtype)
    b = local_quadratic(gb, f_nl).astype(dtype)
    return a, b


def zeldovich_particles(n_side, box, seed, sigma_disp=5.0, amp=P_AMP, k0=P_K0, index=P_INDEX):
    """Lattice particles displaced by the Zel'dovich field of a Gaussian.

    Psi(k) = i k / k^2 delta(k), rescaled so that each component has RMS
    sigma_disp (Mpc/h).  Returns (n_side^3, 3) float32 positions in [0, box).
    """
    dk = gaussian_spectrum(n_side, box, seed, amp, k0, index)
    kf, f, fz = _kgrid(n_side, box)
    kx = kf * f[:, None, None]
    ky = kf * f[None, :, None]
    kz = kf * fz[None, None, :]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    k2[0, 0, 0] = 1.0
    disp = []
    for kc in (kx, ky, kz):
        psi = field_from_spectrum(1j * kc / k2 * dk, n_side)
        disp.append(psi)
    rms = np.sqrt(np.mean([np.mean(p * p) for p in disp]))
    scale = sigma_disp / rms if rms > 0 else 0.0
    q = (np.arange(n_side) + 0.0) * (box / n_side)
    gx, gy, gz = np.meshgrid(q, q, q, indexing='ij')
    pos = np.stack([gx + scale * disp[0], gy + scale * disp[1], gz + scale * disp[2]], axis=-1)
    pos = np.mod(pos, box).reshape(-1, 3).astype(np.float32)
    pos[pos >= box] = 0.0
    return pos


# ---------------------------------------------------------------------------
# Device generators (serving and batch workloads).  Same power spectrum and
# normalisation as above, but the white noise comes from torch's seeded
# device generator (Philox), so the fields differ from the PCG64 ones.  The
# transforms here prepare inputs and are not part of the measured path.
# ---------------------------------------------------------------------------
def _device_kmag(N, box, device):
    import torch
    kf = 2.0 * np.pi / box
    f = torch.fft.fftfreq(N, 1.0 / N, device=device, dtype=torch.float64)
    fz = torch.arange(N // 2 + 1, device=device, dtype=torch.float64)
    return kf, f, fz


def gaussian_field_device(N, box, seed, device='cuda', amp=P_AMP, k0=P_K0, index=P_INDEX):
    """Seeded Gaussian field with P(k) = amp (k/k0)^index, float32 on device."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    w = torch.randn((N, N, N), generator=gen, device=device, dtype=torch.float32)
    wk = torch.fft.rfftn(w)
    del w
    kf, f, fz = _device_kmag(N, box, device)
    vol = float(box) ** 3
    n3 = float(N) ** 3
    for ix in range(N):
        k = kf * torch.sqrt(f[ix] ** 2 + f[:, None] ** 2 + fz[None, :] ** 2)
        p = torch.where(k > 0, amp * (k.clamp_min(1e-30) / k0) ** index, torch.zeros_like(k))
        # r2c / N^3 * sqrt(P N^3 / V), then the unnormalised c2r (irfftn * N^3)
        wk[ix] *= torch.sqrt(p * n3 / vol).to(torch.float32)
    return torch.fft.irfftn(wk, s=(N, N, N))


def local_quadratic_device(d, f_nl=0.5):
    """d + f_nl (d^2 - <d^2>), float32 on device."""
    d2 = d * d
    return d + f_nl * (d2 - d2.mean())


def quadratic_field_device(N, box, seed, device='cuda', f_nl=0.5):
    return local_quadratic_device(gaussian_field_device(N, box, seed, device), f_nl)


def correlated_pair_device(N, box, seed_a, seed_c, device='cuda', rho=0.8, f_nl=0.5):
    """(A, B): A = Q(G_A), B = Q(rho G_A + sqrt(1 - rho^2) G_C) with Q the local
    quadratic transform; rho = 0.8 gives weights 0.8 and 0.6."""
    ga = gaussian_field_device(N, box, seed_a, device)
    gc = gaussian_field_device(N, box, seed_c, device)
    gb = rho * ga + float(np.sqrt(1.0 - rho * rho)) * gc
    del gc
    return local_quadratic_device(ga, f_nl), local_quadratic_device(gb, f_nl)
