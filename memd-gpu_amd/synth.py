"""Seeded synthetic epochs for parity tests and the benchmark (numpy PCG64).

Every generator is a pure function of its arguments and seed, so the same
epoch is produced on any machine with this numpy version.  Shapes follow the
workloads the engine is measured on:

* ``sines_epoch``: n mixed sums of 5/20/60 Hz sinusoids (amplitudes U(0.5,1.5),
  phases U(0,2pi)), mixed by an n x n N(0,1)/sqrt(n) matrix, plus N(0,0.1^2)
  noise; fs = 1000 Hz.  (8 x 1000 host case; 32 x 4096 headline shape.)
* ``eeg_epoch``: per-source 1/f (pink) noise scaled to RMS 30 uV plus 10 Hz
  alpha bursts (Gaussian window, duration U(0.5,2) s, 20 uV) at seeded onsets,
  spatially correlated through the Cholesky factor of exp(-d_ij/lambda) over
  electrodes on a unit sphere; optional 40 Hz 5 uV sinusoid on a random quarter
  of the channels; fs = 512 Hz.
* ``batch_epochs``: uneven epochs, n in {8,16,32,64}, T in [1000, 8000], K = 2n.
"""
import numpy as np


def sines_epoch(n, T, seed, fs=1000.0, freqs=(5.0, 20.0, 60.0), noise=0.1):
    rng = np.random.default_rng(seed)
    t = np.arange(T) / fs
    amp = rng.uniform(0.5, 1.5, (n, len(freqs)))
    phase = rng.uniform(0.0, 2.0 * np.pi, (n, len(freqs)))
    src = np.zeros((n, T))
    for j, f in enumerate(freqs):
        src += amp[:, j:j + 1] * np.sin(2.0 * np.pi * f * t[None, :] + phase[:, j:j + 1])
    mix = rng.standard_normal((n, n)) / np.sqrt(n)
    y = mix @ src + noise * rng.standard_normal((n, T))
    return np.ascontiguousarray(y, np.float64)


def _sphere_points(n):
    # Fibonacci lattice on the unit sphere (deterministic electrode layout)
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = np.pi * (3.0 - np.sqrt(5.0)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def eeg_epoch(n, T, seed, fs=512.0, rms=30.0, alpha_amp=20.0, lam=0.5, gamma40=False):
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, T))
    F = np.fft.rfft(white, axis=1)
    f = np.fft.rfftfreq(T, 1.0 / fs)
    f[0] = f[1] if len(f) > 1 else 1.0
    pink = np.fft.irfft(F / np.sqrt(f)[None, :], n=T, axis=1)
    pink -= pink.mean(axis=1, keepdims=True)
    pink *= rms / np.maximum(pink.std(axis=1, keepdims=True), 1e-30)
    t = np.arange(T) / fs
    dur_total = T / fs
    for c in range(n):
        nb = max(1, int(round(dur_total / 4.0)))
        for _ in range(nb):
            onset = rng.uniform(0.0, dur_total)
            dur = rng.uniform(0.5, 2.0)
            ph = rng.uniform(0.0, 2.0 * np.pi)
            win = np.exp(-0.5 * ((t - onset - dur / 2.0) / (dur / 4.0)) ** 2)
            pink[c] += alpha_amp * win * np.sin(2.0 * np.pi * 10.0 * t + ph)
    pos = _sphere_points(n)
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    C = np.exp(-d / lam) + 1e-9 * np.eye(n)
    Lc = np.linalg.cholesky(C)
    y = Lc @ pink
    if gamma40:
        chans = rng.choice(n, size=max(1, n // 4), replace=False)
        ph = rng.uniform(0.0, 2.0 * np.pi, len(chans))
        y[chans] += 5.0 * np.sin(2.0 * np.pi * 40.0 * t[None, :] + ph[:, None])
    return np.ascontiguousarray(y, np.float64)


def batch_shapes(count, seed, n_choices=(8, 16, 32, 64), t_range=(1000, 8000)):
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.array(n_choices), size=count)
    Ts = rng.integers(t_range[0], t_range[1] + 1, size=count)
    return [(int(n), int(T), int(2 * n)) for n, T in zip(ns, Ts)]


def batch_epochs(count, seed, **kw):
    """[(y, K)] for `count` uneven epochs; epoch i uses seed + 1 + i."""
    out = []
    for i, (n, T, K) in enumerate(batch_shapes(count, seed, **kw)):
        out.append((eeg_epoch(n, T, seed + 1 + i), K))
    return out


def load_eeg_sample(path):
    """The reference's 32 x 30504 float32 channel-major sample file (or its .npz copy)."""
    if path.endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            return np.ascontiguousarray(z["y"])
    raw = np.fromfile(path, dtype="<f4")
    return raw.reshape(32, -1)
