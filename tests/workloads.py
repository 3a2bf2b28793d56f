"""Synthetic inputs of the benchmark configurations, and index maps that cut a
sub-problem (some m-blocks, radii, fields, theta rows) out of a full-size array.

The transforms are independent per column (r, re/im, field) and per m-block, so a
sub-problem cut from a full-size input and pushed through the oracle must equal the
same cut of the full-size device output.  That is how the parity tests check the
configurations whose full size the oracle cannot finish in seconds.

Layouts (the reference's storage):
  spectral  packed over mp in m_values order, block mp = data(m:lmax, r, imi, f), l fastest
  hybrid    p2[mp][c][theta], c = r + imi nr + 2 nr f
  physical  p3[f][theta][r][nphi + 2]
  radial    p1[f][lm][imi][r]   (one Chebyshev column per (f, lm, imi))
"""
from __future__ import annotations

import numpy as np

# the benchmark configurations: (l_max, N_theta, N_phi, N_r, N_f)
CONFIG0 = dict(lmax=63, nt=96, nphi=192, nr=48, nf=4)
CONFIG1 = dict(lmax=255, nt=384, nphi=768, nr=128, nf=8)
CONFIG2 = dict(lmax=511, nt=768, nphi=1536, nr=192, nf=8)
CONFIG4 = dict(lmax=1023, nt=1536, nphi=3072, nr=256, nf=5)
RMIN, RMAX = 0.35, 1.0  # radial domain of the Chebyshev stages (aspect ratio 0.35)


def spectral_offsets(lmax: int, mvals, nr: int, nf: int) -> np.ndarray:
    nl = lmax - np.asarray(mvals, dtype=np.int64) + 1
    return np.concatenate([[0], np.cumsum(nl * 2 * nr * nf)]).astype(np.int64)


def batch_problems(seed: int = 3, n: int = 64):
    """Independent transforms: l_max log-uniform in [31, 1023] with l_max + 1 a multiple
    of 4 (so that N_theta = 3 (l_max + 1) / 2 is even, ProblemSize.F90:443,719), N_r
    from {32, 64, 96}, 1-8 fields; (l_max, N_theta, N_r, N_f) each."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log(32), np.log(1024), n)
    l1 = np.clip(4 * np.round(np.exp(u) / 4), 32, 1024).astype(int)
    nr = rng.choice([32, 64, 96], n)
    nf = rng.integers(1, 9, n)
    return [(int(a - 1), int(3 * a // 2), int(b), int(c)) for a, b, c in zip(l1, nr, nf)]


# ---------------------------------------------------------------- device inputs
def spectral_device(lmax: int, mvals, nr: int, nf: int, seed: int):
    """Coefficients a_lm(r) = (N(0,1) + i N(0,1)) (l+1)^-2 sin(pi r_i / N_r), i = 1..N_r,
    m = 0 real, generated on the device in the packed spectral layout."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    off = spectral_offsets(lmax, mvals, nr, nf)
    s = torch.empty(int(off[-1]), dtype=torch.float64, device="cuda")
    rfac = torch.sin(torch.pi * torch.arange(1, nr + 1, dtype=torch.float64, device="cuda") / nr)
    lf = (torch.arange(0, lmax + 1, dtype=torch.float64, device="cuda") + 1.0) ** -2
    for mp, m in enumerate(mvals):
        m = int(m)
        nl = lmax - m + 1
        blk = torch.randn((nf, 2, nr, nl), dtype=torch.float64, device="cuda", generator=g)
        blk *= lf[m:].view(1, 1, 1, nl)
        blk *= rfac.view(1, 1, nr, 1)
        if m == 0:
            blk[:, 1] = 0.0
        s[int(off[mp]): int(off[mp + 1])] = blk.reshape(-1)
    return s


def cheb_coeffs_device(lmax: int, l_of_lm, m_of_lm, nr: int, nf: int, seed: int):
    """Chebyshev coefficients in the radial layout p1[f][lm][imi][n]: (N(0,1) + i N(0,1))
    (l+1)^-2 (n+1)^-3, m = 0 real."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    nlm = len(l_of_lm)
    c = torch.randn((nf, nlm, 2, nr), dtype=torch.float64, device="cuda", generator=g)
    lf = (torch.as_tensor(np.asarray(l_of_lm), dtype=torch.float64, device="cuda") + 1.0) ** -2
    nfac = (torch.arange(nr, dtype=torch.float64, device="cuda") + 1.0) ** -3
    c *= lf.view(1, nlm, 1, 1)
    c *= nfac.view(1, 1, 1, nr)
    m0 = torch.as_tensor(np.asarray(m_of_lm) == 0, device="cuda")
    c[:, m0, 1, :] = 0.0
    return c.reshape(-1)


# ---------------------------------------------------------------- sub-problem maps
def spectral_index(lmax: int, mvals, nr: int, nf: int, mps, rs=None, fs=None) -> np.ndarray:
    """Flat indices into a packed spectral array of the sub-problem (m-blocks mps, radii
    rs, fields fs), in the sub-problem's own packed order."""
    rs = np.arange(nr) if rs is None else np.asarray(rs)
    fs = np.arange(nf) if fs is None else np.asarray(fs)
    off = spectral_offsets(lmax, mvals, nr, nf)
    parts = []
    for mp in mps:
        nl = lmax - int(mvals[mp]) + 1
        f, i, r, l = np.ix_(fs, np.arange(2), rs, np.arange(nl))
        parts.append((off[mp] + ((f * 2 + i) * nr + r) * nl + l).ravel())
    return np.concatenate(parts).astype(np.int64)


def hybrid_index(nt: int, nr: int, nf: int, mps, rs=None, fs=None, thetas=None) -> np.ndarray:
    """Flat indices into a hybrid array p2[mp][c][theta] of the sub-problem, in the
    sub-problem's order [mp'][c'][theta'] with c' = r' + imi nr' + 2 nr' f'."""
    rs = np.arange(nr) if rs is None else np.asarray(rs)
    fs = np.arange(nf) if fs is None else np.asarray(fs)
    th = np.arange(nt) if thetas is None else np.asarray(thetas)
    nrhs = 2 * nr * nf
    mp, f, i, r, t = np.ix_(np.asarray(mps), fs, np.arange(2), rs, th)
    return ((mp * nrhs + r + i * nr + 2 * nr * f) * nt + t).ravel().astype(np.int64)


def grid_index(nt: int, nr: int, nf: int, nphi: int, rs=None, fs=None, thetas=None) -> np.ndarray:
    """Flat indices into a physical array p3[f][theta][r][nphi + 2] of the sub-problem
    (whole lines), in its order [f'][theta'][r'][q]."""
    rs = np.arange(nr) if rs is None else np.asarray(rs)
    fs = np.arange(nf) if fs is None else np.asarray(fs)
    th = np.arange(nt) if thetas is None else np.asarray(thetas)
    f, t, r, q = np.ix_(fs, th, rs, np.arange(nphi + 2))
    return (((f * nt + t) * nr + r) * (nphi + 2) + q).ravel().astype(np.int64)


def radial_index(nlm: int, nr: int, nf: int, rs=None, fs=None) -> np.ndarray:
    """Flat indices into a radial array p1[f][lm][imi][r] of the sub-problem (radii rs,
    fields fs, every (l, m) pair), in its order [f'][lm][imi][r']."""
    rs = np.arange(nr) if rs is None else np.asarray(rs)
    fs = np.arange(nf) if fs is None else np.asarray(fs)
    f, lm, i, r = np.ix_(fs, np.arange(nlm), np.arange(2), rs)
    return (((f * nlm + lm) * 2 + i) * nr + r).ravel().astype(np.int64)


def column_index(n: int, cols) -> np.ndarray:
    """Flat indices of whole radial columns (n values each) of a [col][n] buffer."""
    cols = np.asarray(cols, dtype=np.int64)
    return (cols[:, None] * n + np.arange(n)[None, :]).ravel()


def gather(t, idx: np.ndarray) -> np.ndarray:
    """t[idx] for a device tensor t, copied to the host."""
    import torch
    i = torch.from_numpy(idx).to(t.device)
    return t[i].cpu().numpy()


def rel_max(a, b) -> float:
    """max |a - b| / max |b|"""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_max_dev(a, b) -> float:
    """max |a - b| / max |b| of two device tensors."""
    return float(((a - b).abs().max() / b.abs().max().clamp_min(1e-300)).item())
